// pbf_adjoint_ref.cpp -- TEST INFRASTRUCTURE ONLY: the checker of the GPU
// adjoint.  Nothing in the library links or calls this.
//
// A restatement of the substep of pbf_oracle.c (which follows
// assets/kernels/*.cl, driven as src/Simulation.cpp:554-621) with the integer
// structure frozen -- the sort permutation, the friend lists and the wall
// planes that fired, all recorded from the binary32 oracle -- written once
// over a scalar type T:
//   T = double                 the FP64 function, for central finite differences;
//   T = Dual<double>           its forward-mode directional derivative (JVP);
//   T = Dual<float>            binary32 values with FP64 tangents: the exact
//                              derivative of the binary32 trajectory the GPU
//                              computes (its values equal the oracle's bit for
//                              bit, because every operation is in the oracle's
//                              order, fused by t_fma where the oracle calls
//                              fmaf, and the build does no other contraction).
// With the structure frozen the function is smooth: every kernel is C1 at
// the r = h cutoff, and the remaining branches (the density mask of XSPH and
// vorticity, eta = 0) are taken on the primal values.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

extern "C" {
#include "pbf_oracle.h"
}

namespace {

template <class P>
struct Dual {
    P v;
    double d;
    Dual() : v(0), d(0.0) {}
    Dual(P value, double tangent = 0.0) : v(value), d(tangent) {}
};

template <class P> inline Dual<P> operator+(Dual<P> a, Dual<P> b) { return {a.v + b.v, a.d + b.d}; }
template <class P> inline Dual<P> operator-(Dual<P> a, Dual<P> b) { return {a.v - b.v, a.d - b.d}; }
template <class P> inline Dual<P> operator-(Dual<P> a) { return {-a.v, -a.d}; }
template <class P> inline Dual<P> operator*(Dual<P> a, Dual<P> b) {
    return {a.v * b.v, a.d * (double)b.v + (double)a.v * b.d};
}
template <class P> inline Dual<P> operator/(Dual<P> a, Dual<P> b) {
    const double bv = (double)b.v;
    return {a.v / b.v, (a.d * bv - (double)a.v * b.d) / (bv * bv)};
}
template <class P> inline Dual<P>& operator+=(Dual<P>& a, Dual<P> b) { return a = a + b; }
template <class P> inline Dual<P>& operator*=(Dual<P>& a, Dual<P> b) { return a = a * b; }
template <class P> inline Dual<P>& operator/=(Dual<P>& a, Dual<P> b) { return a = a / b; }

inline double val(double x) { return x; }
inline double val(float x) { return (double)x; }
template <class P> inline double val(Dual<P> x) { return (double)x.v; }

inline double t_sqrt(double x) { return sqrt(x); }
inline float t_sqrt(float x) { return sqrtf(x); }
inline Dual<double> t_sqrt(Dual<double> x) {
    const double s = sqrt(x.v);
    return {s, x.d / (2.0 * s)};
}
inline Dual<float> t_sqrt(Dual<float> x) {
    const float s = sqrtf(x.v);
    return {s, x.d / (2.0 * (double)s)};
}

// The oracle's fmaf at the same points: one rounding in the value's own
// precision; a dual's tangent is that of a * b + c.
inline double t_fma(double a, double b, double c) { return fma(a, b, c); }
inline float t_fma(float a, float b, float c) { return fmaf(a, b, c); }
template <class P> inline Dual<P> t_fma(Dual<P> a, Dual<P> b, Dual<P> c) {
    return {t_fma(a.v, b.v, c.v), a.d * (double)b.v + (double)a.v * b.d + c.d};
}
// r^2 of a pair as the oracle's dist2
template <class T> inline T dist2(const T* r) { return t_fma(r[2], r[2], t_fma(r[1], r[1], r[0] * r[0])); }

// Wall noise: sin in FP64; the binary32 oracle's own sine for Dual<float>.
inline double t_sin(double x) { return sin(x); }
inline float t_sin(float x) { return pbfo_sinf(x); }
inline Dual<double> t_sin(Dual<double> x) { return {sin(x.v), x.d * cos(x.v)}; }
inline Dual<float> t_sin(Dual<float> x) { return {pbfo_sinf(x.v), x.d * cos((double)x.v)}; }

// Scalar constant of the computation type (binary32 constants for Dual<float>).
template <class T> struct K;
template <> struct K<double> { static double c(float f) { return (double)f; } };
template <> struct K<float> { static float c(float f) { return f; } };
template <> struct K<Dual<double>> { static Dual<double> c(float f) { return {(double)f, 0.0}; } };
template <> struct K<Dual<float>> { static Dual<float> c(float f) { return {f, 0.0}; } };

// |r| from r^2 as the oracle's sweeps take it: 0 below 2^-96 (a coincident
// pair, no direction), else the square root.
template <class T>
inline T pair_rl(T r2) {
    return val(r2) >= 0x1p-96 ? t_sqrt(r2) : K<T>::c(0.0f);
}

struct Frozen {
    uint32_t n, iters, steps;
    const uint32_t* perm;      // steps x n
    const uint32_t* friends;   // steps x (5 + 250) n, reference layout
    const uint32_t* hits;      // steps x iters x n: planes that moved each particle
};

inline uint32_t fr_count(const uint32_t* fr, uint32_t n, int q, uint32_t i) { return fr[(size_t)q * n + i]; }
inline uint32_t fr_idx(const uint32_t* fr, uint32_t n, int q, int f, uint32_t i) {
    return fr[(size_t)PBFO_CIRCLES * n + ((size_t)q * PBFO_PER_CIRCLE + f) * n + i];
}

// Friends a sweep with threshold th processes (compute_scaling.cl:50,
// compute_delta.cl:100): calls fn(j) in the reference's order.
template <class Fn>
void for_friends(const uint32_t* fr, uint32_t n, uint32_t i, float th, Fn&& fn) {
    int cnt[PBFO_CIRCLES], total = 0;
    for (int q = 0; q < PBFO_CIRCLES; ++q) total += cnt[q] = (int)fr_count(fr, n, q, i);
    int done = 0;
    for (int q = 0; q < PBFO_CIRCLES; ++q) {
        if (pbfo_skip_circle(done, total, th)) continue;
        done += cnt[q];
        for (int f = 0; f < cnt[q]; ++f) fn(fr_idx(fr, n, q, f, i));
    }
}

template <class T>
struct V {
    T a[3];
};

// State of a rollout in the sorted order of the last substep; ids[i] is the
// identity of the particle at slot i.
template <class T>
struct State {
    std::vector<V<T>> x, v;
    std::vector<uint32_t> ids;
};

template <class T>
State<T> state_from(uint32_t n, const T* x0, const T* v0) {
    State<T> st;
    st.x.resize(n);
    st.v.resize(n);
    st.ids.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            st.x[i].a[k] = x0[3 * i + k];
            st.v[i].a[k] = v0[3 * i + k];
        }
        st.ids[i] = i;
    }
    return st;
}

template <class T>
void state_to(const State<T>& st, T* xT, T* vT) {
    for (size_t i = 0; i < st.ids.size(); ++i)
        for (int k = 0; k < 3; ++k) {
            if (xT) xT[3 * st.ids[i] + k] = st.x[i].a[k];
            if (vT) vT[3 * st.ids[i] + k] = st.v[i].a[k];
        }
}

// Substep s of the frozen structure applied to st.  coef = {viscosity,
// vorticity, g0, g1, g2}.
template <class T>
void substep(const pbfo_cfg& c, const Frozen& f, uint32_t s, const T* coef, State<T>& st) {
    const uint32_t n = f.n;
    const float hf = c.h, h2f = hf * hf;
    float poly6f, spikyf;
    pbfo_kernel_factors(hf, &poly6f, &spikyf);
    const float dqh = c.st_dist * hf;
    const float q2 = dqh * dqh;
    const float tq = h2f - q2;
    const float poly6_qf = tq * tq * tq;
    const T h = K<T>::c(hf), h2 = K<T>::c(h2f), dt = K<T>::c(c.dt), rd = K<T>::c(c.rest_density);
    const T eps = K<T>::c(c.epsilon), stk = K<T>::c(c.st_k), poly6 = K<T>::c(poly6f), spiky = K<T>::c(spikyf);
    const T inv_q = K<T>::c(1.0f / poly6_qf);
    const T visc = coef[0], vort = coef[1];
    const T g[3] = {coef[2], coef[3], coef[4]};
    const T zero = K<T>::c(0.0f), one = K<T>::c(1.0f), five = K<T>::c(5.0f), three = K<T>::c(3.0f);
    const T c003 = K<T>::c(0.03f), minus1 = K<T>::c(-1.0f);

    // collision quads with the constants BouncePointQuad derives (binary32,
    // as the oracle computes them)
    struct QuadF {
        float b[3], e0[3], e1[3], n[3], aa, bb, cc, invdet;
    };
    const uint32_t qn = c.quads && c.mask ? (c.quad_count < 26 ? c.quad_count : 26) : 0;
    std::vector<QuadF> qs(qn);
    for (uint32_t k = 0; k < qn; ++k) {
        const pbfo_quad& q = c.quads[k];
        QuadF& o = qs[k];
        for (int a = 0; a < 3; ++a) { o.b[a] = q.b[a]; o.e0[a] = q.e0[a]; o.e1[a] = q.e1[a]; }
        o.aa = q.e0[0] * q.e0[0] + q.e0[1] * q.e0[1] + q.e0[2] * q.e0[2];
        o.bb = q.e0[0] * q.e1[0] + q.e0[1] * q.e1[1] + q.e0[2] * q.e1[2];
        o.cc = q.e1[0] * q.e1[0] + q.e1[1] * q.e1[1] + q.e1[2] * q.e1[2];
        o.invdet = 1.0f / (o.aa * o.cc - o.bb * o.bb);
        const float cr[3] = {q.e0[1] * q.e1[2] - q.e0[2] * q.e1[1], q.e0[2] * q.e1[0] - q.e0[0] * q.e1[2],
                             q.e0[0] * q.e1[1] - q.e0[1] * q.e1[0]};
        const float il = 1.0f / sqrtf(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        for (int a = 0; a < 3; ++a) o.n[a] = cr[a] * il;
    }
    std::vector<V<T>>& x = st.x;
    std::vector<V<T>>& v = st.v;
    std::vector<V<T>> xs(n), xs2(n), tmp(n), om(n);
    std::vector<T> lam(n), rho(n);
    std::vector<uint32_t> ids2(n);
    {
        const uint32_t* perm = f.perm + (size_t)s * n;
        const uint32_t* fr = f.friends + (size_t)s * (PBFO_CIRCLES + PBFO_CIRCLES * PBFO_PER_CIRCLE) * n;
        // predict_positions.cl (v in the previous order), then the reorder
        for (uint32_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) {
                v[i].a[k] = v[i].a[k] + dt * g[k];
                tmp[i].a[k] = x[i].a[k] + dt * v[i].a[k];
            }
        for (uint32_t i = 0; i < n; ++i) {
            xs[i] = tmp[perm[i]];
            ids2[i] = st.ids[perm[i]];
        }
        for (uint32_t i = 0; i < n; ++i) tmp[i] = x[perm[i]];
        x.swap(tmp);
        st.ids.swap(ids2);
        for (uint32_t it = 0; it < f.iters; ++it) {
            const uint32_t* hit = f.hits + ((size_t)s * f.iters + it) * n;
            // compute_scaling.cl
#pragma omp parallel for schedule(static)
            for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
                const uint32_t i = (uint32_t)ii;
                const T* pi = xs[i].a;
                T dens = zero, gsum = zero, gi[3] = {zero, zero, zero};
                for_friends(fr, n, i, 0.6f, [&](uint32_t j) {
                    const T* pj = xs[j].a;
                    const T r[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
                    const T r2 = dist2(r);
                    if (val(r2) < val(h2)) {
                        const T rl = pair_rl(r2);
                        const T hr = h - rl;
                        const T sp = spiky * hr * hr;
                        const T inv = val(rl) > 0.0 ? one / rl : zero;
                        const T t = sp * inv;
                        const T gg[3] = {r[0] * t, r[1] * t, r[2] * t};
                        const T d = h2 - r2;
                        dens = t_fma(d * d, d, dens);
                        gsum = t_fma(gg[2], gg[2], t_fma(gg[1], gg[1], t_fma(gg[0], gg[0], gsum)));
                        gi[0] += gg[0]; gi[1] += gg[1]; gi[2] += gg[2];
                    }
                });
                dens *= poly6;
                rho[i] = dens;
                gsum += gi[0] * gi[0] + gi[1] * gi[1] + gi[2] * gi[2];
                const T C = (dens / rd) - one;
                lam[i] = minus1 * C / (gsum / (rd * rd) + eps);
            }
            // compute_delta.cl with the recorded walls, then update_predicted.cl
#pragma omp parallel for schedule(static)
            for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
                const uint32_t i = (uint32_t)ii;
                const T* pi = xs[i].a;
                T sum[3] = {zero, zero, zero};
                for_friends(fr, n, i, 0.5f, [&](uint32_t j) {
                    const T* pj = xs[j].a;
                    const T r[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
                    const T r2 = dist2(r);
                    if (val(r2) < val(h2)) {
                        const T rl = pair_rl(r2);
                        const T hr = h - rl;
                        const T inv = val(rl) > 0.0 ? one / rl : zero;
                        const T t = inv * hr * hr;
                        const T gg[3] = {r[0] * t, r[1] * t, r[2] * t};
                        const T d = h2 - r2;
                        const T p6 = d * d * d;
                        const T ratio = p6 * inv_q;
                        const T r4 = ratio * ratio;
                        const T co = t_fma(stk, r4 * r4, lam[i] + lam[j]);
                        for (int k = 0; k < 3; ++k) sum[k] = t_fma(co, gg[k], sum[k]);
                    }
                });
                T fut[3];
                for (int k = 0; k < 3; ++k) fut[k] = pi[k] + (minus1 * spiky * sum[k]) / rd;
                if (hit[i]) {
                    const T ns[3] = {t_sin(fut[0] * five / h), t_sin(fut[1] * five / h), t_sin(fut[2] * five / h)};
                    const T edge = (three + ns[0] + ns[1] + ns[2]) * h * c003;
                    for (int a = 0; a < 3; ++a) {
                        if (hit[i] & (1u << (2 * a))) fut[a] = K<T>::c(c.box_lo[a]) + edge;
                        if (hit[i] & (2u << (2 * a))) fut[a] = K<T>::c(c.box_hi[a]) - edge;
                    }
                    // quads (BouncePointQuad): the projection onto the plane, offset
                    const T eq = K<T>::c(c.edge_base) + edge;
                    for (uint32_t k = 0; k < qn; ++k) {
                        if (!(hit[i] & (1u << (6 + k)))) continue;
                        const QuadF& q = qs[k];
                        const T D[3] = {K<T>::c(q.b[0]) - fut[0], K<T>::c(q.b[1]) - fut[1], K<T>::c(q.b[2]) - fut[2]};
                        const T e = K<T>::c(q.e1[0]) * D[0] + K<T>::c(q.e1[1]) * D[1] + K<T>::c(q.e1[2]) * D[2];
                        const T d = K<T>::c(q.e0[0]) * D[0] + K<T>::c(q.e0[1]) * D[1] + K<T>::c(q.e0[2]) * D[2];
                        const T ss = K<T>::c(q.invdet) * (K<T>::c(q.bb) * e - K<T>::c(q.cc) * d);
                        const T tt = K<T>::c(q.invdet) * (K<T>::c(q.bb) * d - K<T>::c(q.aa) * e);
                        for (int a = 0; a < 3; ++a)
                            fut[a] = K<T>::c(q.b[a]) + K<T>::c(q.e0[a]) * ss + K<T>::c(q.e1[a]) * tt + K<T>::c(q.n[a]) * eq;
                    }
                }
                for (int k = 0; k < 3; ++k) xs2[i].a[k] = pi[k] + (fut[k] - pi[k]);
            }
            xs.swap(xs2);
        }
        // update_velocities.cl
        for (uint32_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) {
                v[i].a[k] = (xs[i].a[k] - x[i].a[k]) / dt;
                x[i].a[k] = xs[i].a[k];
            }
        // apply_viscosity.cl (Jacobi), omega
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
            const uint32_t i = (uint32_t)ii;
            const T* pi = xs[i].a;
            const T* vi = v[i].a;
            T vs[3] = {zero, zero, zero}, o[3] = {zero, zero, zero};
            for_friends(fr, n, i, 0.5f, [&](uint32_t j) {
                const T* pj = xs[j].a;
                const T* vj = v[j].a;
                const T r[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
                const T r2 = dist2(r);
                if (val(r2) < val(h2) && fabs(val(rho[j])) > (double)1e-8f) {
                    const T u[3] = {vj[0] - vi[0], vj[1] - vi[1], vj[2] - vi[2]};
                    const T d = h2 - r2;
                    const T rl = pair_rl(r2);
                    const T hr = h - rl;
                    const T inv = val(rl) > 0.0 ? one / rl : zero;
                        const T t = inv * hr * hr;
                        const T gg[3] = {r[0] * t, r[1] * t, r[2] * t};
                    o[0] += t_fma(u[1], gg[2], -(u[2] * gg[1]));
                    o[1] += t_fma(u[2], gg[0], -(u[0] * gg[2]));
                    o[2] += t_fma(u[0], gg[1], -(u[1] * gg[0]));
                    const T p6 = d * d * d;
                    for (int k = 0; k < 3; ++k) vs[k] = t_fma(u[k], p6, vs[k]);
                }
            });
            for (int k = 0; k < 3; ++k) {
                tmp[i].a[k] = vi[k] + visc * (vs[k] * poly6);
                om[i].a[k] = o[k] * (minus1 * spiky);
            }
        }
        v.swap(tmp);
        // apply_vorticity.cl
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
            const uint32_t i = (uint32_t)ii;
            const T* pi = xs[i].a;
            T eta[3] = {zero, zero, zero};
            for_friends(fr, n, i, 0.5f, [&](uint32_t j) {
                const T* pj = xs[j].a;
                const T r[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
                const T r2 = dist2(r);
                if (val(r2) < val(h2) && fabs(val(rho[j])) > (double)1e-8f) {
                    const T rl = pair_rl(r2);
                    const T hr = h - rl;
                    const T inv = val(rl) > 0.0 ? one / rl : zero;
                        const T t = inv * hr * hr;
                        const T gg[3] = {r[0] * t, r[1] * t, r[2] * t};
                    const T* oj = om[j].a;
                    const T wl = pair_rl(oj[0] * oj[0] + oj[1] * oj[1] + oj[2] * oj[2]);
                    for (int k = 0; k < 3; ++k) eta[k] = t_fma(wl, gg[k], eta[k]);
                }
            });
            for (int k = 0; k < 3; ++k) eta[k] *= minus1 * spiky;
            const T l = t_sqrt(eta[0] * eta[0] + eta[1] * eta[1] + eta[2] * eta[2]);
            if (val(l) > 0.0) { eta[0] /= l; eta[1] /= l; eta[2] /= l; }
            const T* oi = om[i].a;
            const T cr[3] = {eta[1] * oi[2] - eta[2] * oi[1], eta[2] * oi[0] - eta[0] * oi[2],
                             eta[0] * oi[1] - eta[1] * oi[0]};
            for (int k = 0; k < 3; ++k) v[i].a[k] += (vort * cr[k]) * dt;
        }
    }
}

// A rollout of f.steps substeps; positions/velocities in identity order.
template <class T>
void rollout(const pbfo_cfg& c, const Frozen& f, const T* x0, const T* v0, const T* coef, T* xT, T* vT) {
    State<T> st = state_from<T>(f.n, x0, v0);
    for (uint32_t s = 0; s < f.steps; ++s) substep<T>(c, f, s, coef, st);
    state_to<T>(st, xT, vT);
}

template <class P>
void jvp(const pbfo_cfg& c, const Frozen& f, const P* x0, const P* v0, const P* coef, const double* dx0,
         const double* dv0, const double* dcoef, P* xT, P* vT, double* dxT, double* dvT) {
    using D = Dual<P>;
    const size_t m = 3 * (size_t)f.n;
    std::vector<D> X(m), Vv(m), XT(m), VT(m);
    D C[5];
    for (size_t i = 0; i < m; ++i) {
        X[i] = D(x0[i], dx0 ? dx0[i] : 0.0);
        Vv[i] = D(v0[i], dv0 ? dv0[i] : 0.0);
    }
    for (int k = 0; k < 5; ++k) C[k] = D(coef[k], dcoef ? dcoef[k] : 0.0);
    rollout<D>(c, f, X.data(), Vv.data(), C, XT.data(), VT.data());
    for (size_t i = 0; i < m; ++i) {
        if (xT) xT[i] = XT[i].v;
        if (vT) vT[i] = VT[i].v;
        if (dxT) dxT[i] = XT[i].d;
        if (dvT) dvT[i] = VT[i].d;
    }
}

// FP64 central differences of every substep map, chained along the binary32
// trajectory: the tangent t_s is pushed through substep s by
// (F64_s(S32_s + e t_s) - F64_s(S32_s - e t_s)) / 2e at the binary32 state
// S32_s, with e scaled so that no state component moves by more than
// `delta` and no coefficient by more than `delta` times its size.
void tangent_fd(const pbfo_cfg& c, const Frozen& f, const float* x0, const float* v0, const float* coef,
                const double* dx0, const double* dv0, const double* dcoef, double delta, double* dxT, double* dvT) {
    const uint32_t n = f.n;
    State<float> s32 = state_from<float>(n, x0, v0);
    std::vector<V<double>> tx(n), tv(n);
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            tx[i].a[k] = dx0 ? dx0[3 * i + k] : 0.0;
            tv[i].a[k] = dv0 ? dv0[3 * i + k] : 0.0;
        }
    double tc[5];
    for (int k = 0; k < 5; ++k) tc[k] = dcoef ? dcoef[k] : 0.0;
    for (uint32_t s = 0; s < f.steps; ++s) {
        double tmax = 0.0;
        for (uint32_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) tmax = fmax(tmax, fmax(fabs(tx[i].a[k]), fabs(tv[i].a[k])));
        // a coefficient moves by at most `delta` relative to its own size
        for (int k = 0; k < 5; ++k) tmax = fmax(tmax, fabs(tc[k]) / (coef[k] != 0.0f ? fabs((double)coef[k]) : 1.0));
        const double e = tmax > 0.0 ? delta / tmax : delta;
        State<double> sp, sm;
        for (State<double>* q : {&sp, &sm}) {
            const double sg = q == &sp ? e : -e;
            q->x.resize(n);
            q->v.resize(n);
            q->ids = s32.ids;
            for (uint32_t i = 0; i < n; ++i)
                for (int k = 0; k < 3; ++k) {
                    q->x[i].a[k] = (double)s32.x[i].a[k] + sg * tx[i].a[k];
                    q->v[i].a[k] = (double)s32.v[i].a[k] + sg * tv[i].a[k];
                }
        }
        double cp[5], cm[5];
        for (int k = 0; k < 5; ++k) {
            cp[k] = (double)coef[k] + e * tc[k];
            cm[k] = (double)coef[k] - e * tc[k];
        }
        substep<double>(c, f, s, cp, sp);
        substep<double>(c, f, s, cm, sm);
        substep<float>(c, f, s, coef, s32);
        for (uint32_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) {
                tx[i].a[k] = (sp.x[i].a[k] - sm.x[i].a[k]) / (2.0 * e);
                tv[i].a[k] = (sp.v[i].a[k] - sm.v[i].a[k]) / (2.0 * e);
            }
    }
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            if (dxT) dxT[3 * s32.ids[i] + k] = tx[i].a[k];
            if (dvT) dvT[3 * s32.ids[i] + k] = tv[i].a[k];
        }
}

}  // namespace

extern "C" {

void pbfa_tangent_fd(const pbfo_cfg* c, uint32_t steps, const uint32_t* perm, const uint32_t* friends,
                     const uint32_t* hits, const float* x0, const float* v0, const float* coef, const double* dx0,
                     const double* dv0, const double* dcoef, double delta, double* dxT, double* dvT) {
    const Frozen f{c->n, c->iters, steps, perm, friends, hits};
    tangent_fd(*c, f, x0, v0, coef, dx0, dv0, dcoef, delta, dxT, dvT);
}


// Structure arrays: perm[steps][n], friends[steps][(5 + 250) n] (the
// reference layout of pbfo_buffer(.., 9)), hits[steps][iters][n].
void pbfa_rollout_f64(const pbfo_cfg* c, uint32_t steps, const uint32_t* perm, const uint32_t* friends,
                      const uint32_t* hits, const double* x0, const double* v0, const double* coef, double* xT,
                      double* vT) {
    const Frozen f{c->n, c->iters, steps, perm, friends, hits};
    rollout<double>(*c, f, x0, v0, coef, xT, vT);
}

void pbfa_jvp_f64(const pbfo_cfg* c, uint32_t steps, const uint32_t* perm, const uint32_t* friends,
                  const uint32_t* hits, const double* x0, const double* v0, const double* coef, const double* dx0,
                  const double* dv0, const double* dcoef, double* xT, double* vT, double* dxT, double* dvT) {
    const Frozen f{c->n, c->iters, steps, perm, friends, hits};
    jvp<double>(*c, f, x0, v0, coef, dx0, dv0, dcoef, xT, vT, dxT, dvT);
}

void pbfa_jvp_f32(const pbfo_cfg* c, uint32_t steps, const uint32_t* perm, const uint32_t* friends,
                  const uint32_t* hits, const float* x0, const float* v0, const float* coef, const double* dx0,
                  const double* dv0, const double* dcoef, float* xT, float* vT, double* dxT, double* dvT) {
    const Frozen f{c->n, c->iters, steps, perm, friends, hits};
    jvp<float>(*c, f, x0, v0, coef, dx0, dv0, dcoef, xT, vT, dxT, dvT);
}

}  // extern "C"
