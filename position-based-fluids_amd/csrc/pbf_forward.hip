// pbf_forward.hip -- the per-substep kernels other than the sort.
//
// Kernel chain of one substep (reference: src/Simulation.cpp:554-621):
//   k_predict_hash   predictPositions + computeKeys + digit histograms for the
//                    sort + reset of the previous step's cells
//   (onesweep passes, pbf_sort.hip)
//   k_gather         sortParticles; recomputes x* from the gathered x, v
//   k_cells          updateCells
//   k_friends        buildFriendsList
//   K x { k_lambda   computeScaling + packData(lambda)
//         k_delta    computeDelta (+ walls) + updatePredicted; on the last
//                    iteration also packData(density) + updateVelocities }
//   k_viscosity      applyViscosity (Jacobi: reads a snapshot of v)
//   k_vorticity      applyVorticity
// Particles live in cell-sorted order as float4 arrays; the friends list
// keeps the reference layout (counts packed 5 x 6 bits per particle,
// indices[5][50][n]) so a wave reads one friend slot of 64 consecutive
// particles as one contiguous 256-byte segment.
#include "pbf_internal.h"

namespace pbf {

namespace {

constexpr int kThreads = 256;
constexpr int kPredictItems = 4;

__device__ __forceinline__ float4 predicted_from(const float4 x, const float4 v, const StepConsts& c, uint32_t pause) {
    // predict_positions.cl:12-17
    float4 vg = v;
    if (pause == 0u) {
        vg.x = v.x + c.dt * c.g[0];
        vg.y = v.y + c.dt * c.g[1];
        vg.z = v.z + c.dt * c.g[2];
    }
    float4 o;
    o.x = x.x + c.dt * vg.x;
    o.y = x.y + c.dt * vg.y;
    o.z = x.z + c.dt * vg.z;
    o.w = x.w + c.dt * vg.w;
    return o;
}

__device__ __forceinline__ uint32_t key_of(const float4 p, const StepConsts& c) {
    // radixsort.cl:11-13
    return cell_hash(cell_coord(p.x / c.h), cell_coord(p.y / c.h), cell_coord(p.z / c.h), c);
}

__global__ __launch_bounds__(kThreads) void k_predict_hash(const float4* __restrict__ x, const float4* __restrict__ v,
                                                          uint32_t* __restrict__ keys,
                                                          const uint32_t* __restrict__ old_keys,
                                                          uint2* __restrict__ cells, uint32_t reset_old,
                                                          uint32_t* __restrict__ hist, uint32_t passes,
                                                          uint32_t pause, StepConsts c, uint32_t* err) {
    __shared__ uint32_t s_hist[kMaxPasses * kRadix];
    for (int q = threadIdx.x; q < (int)passes * kRadix; q += kThreads) s_hist[q] = 0;
    __syncthreads();
    bool bad = false;
    const uint32_t base = blockIdx.x * (uint32_t)(kThreads * kPredictItems);
#pragma unroll
    for (int k = 0; k < kPredictItems; ++k) {
        const uint32_t i = base + (uint32_t)(k * kThreads) + threadIdx.x;
        if (i < c.n) {
            const float4 p = predicted_from(x[i], v[i], c, pause);
            const uint32_t key = key_of(p, c);
            keys[i] = key;
            for (uint32_t q = 0; q < passes; ++q) atomicAdd(&s_hist[q * kRadix + ((key >> (kRadixBits * q)) & (kRadix - 1))], 1u);
            if (reset_old) {
                // reset_grid.cl: empty the cells of the previous step
                const uint32_t ok = old_keys[i];
                if ((i == 0 || old_keys[i - 1] != ok) && ok < c.gbs) cells[ok] = make_uint2(kEmpty, kEmpty);
            }
            bad |= !(fabsf(p.x) <= 3.4e38f && fabsf(p.y) <= 3.4e38f && fabsf(p.z) <= 3.4e38f);
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < (int)passes * kRadix; q += kThreads) {
        const uint32_t h = s_hist[q];
        if (h) atomicAdd(&hist[q], h);
    }
    if (bad) atomicOr(err, (uint32_t)PBF_FLAG_NAN);
}

// Paused step with a valid sort: predictPositions only (no re-sort).
__global__ __launch_bounds__(kThreads) void k_predict_only(const float4* __restrict__ x, const float4* __restrict__ v,
                                                          float4* __restrict__ xs, uint32_t pause, StepConsts c) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= c.n) return;
    xs[i] = predicted_from(x[i], v[i], c, pause);
}

// radixsort.cl:22-37.  x* is recomputed from the gathered x and v with the
// same operations as k_predict_hash, so it is bitwise the value that was
// hashed; v itself is not permuted (it is rewritten before it is read).
__global__ __launch_bounds__(kThreads) void k_gather(const uint32_t* __restrict__ perm, const float4* __restrict__ x,
                                                    const float4* __restrict__ v, float4* __restrict__ x_out,
                                                    float4* __restrict__ xs_out, uint32_t pause, StepConsts c,
                                                    uint32_t* err) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= c.n) return;
    uint32_t p = perm[i];
    if (p >= c.n) {  // only after a failed sort (flagged there)
        atomicOr(err, (uint32_t)PBF_FLAG_INTERNAL);
        p = i;
    }
    const float4 xp = x[p];
    x_out[i] = xp;
    xs_out[i] = predicted_from(xp, v[p], c, pause);
}

// update_cells.cl:1-26 (the last particle always closes its cell).
__global__ __launch_bounds__(kThreads) void k_cells(const uint32_t* __restrict__ ks, uint32_t* __restrict__ cells,
                                                   uint32_t n, uint32_t gbs, uint32_t* err) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = ks[i];
    if (k >= gbs || (i > 0 && ks[i - 1] >= gbs)) {  // only after a failed sort
        atomicOr(err, (uint32_t)PBF_FLAG_INTERNAL);
        return;
    }
    if (i == 0) {
        cells[2 * (size_t)k] = 0;
    } else {
        const uint32_t pk = ks[i - 1];
        if (pk != k) {
            cells[2 * (size_t)k] = i;
            cells[2 * (size_t)pk + 1] = i - 1;
        }
    }
    if (i == n - 1) cells[2 * (size_t)k + 1] = i;
}

// build_friends_list.cl:23-103.  27 cells in x, y, z order, ascending j;
// circle = min(int(a^3 * 5), 4), a = max(0, (r - 0.3h) / 0.7h); at most 50
// friends per circle, later ones dropped (and counted).
__global__ __launch_bounds__(kThreads) void k_friends(const float4* __restrict__ xs, const uint2* __restrict__ cells,
                                                     uint32_t* __restrict__ nb_cnt, uint32_t* __restrict__ nb_idx,
                                                     StepConsts c, uint32_t* err) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= c.n) return;
    const float4 p = xs[i];
    const int32_t cx = cell_coord(p.x / c.h), cy = cell_coord(p.y / c.h), cz = cell_coord(p.z / c.h);
    uint32_t packed = 0, dropped = 0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                const uint2 b = cells[cell_hash(cx + dx, cy + dy, cz + dz, c)];
                if (b.x == kEmpty) continue;
                const uint32_t e = min(b.y, c.n - 1);
                for (uint32_t j = b.x; j <= e; ++j) {
                    if (j == i) continue;
                    const float4 q = xs[j];
                    const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
                    const float r2 = rx * rx + ry * ry + rz * rz;
                    if (r2 >= c.h2) continue;
                    const float a = fmaxf(0.0f, (sqrtf(r2) - c.min_r) / c.span);
                    const int circ = min((int)(a * a * a * (float)kCircles), kCircles - 1);
                    const uint32_t cnt = (packed >> (6 * circ)) & 63u;
                    if (cnt >= (uint32_t)kPerCircle) { ++dropped; continue; }
                    nb_idx[friend_slot(circ, (int)cnt, i, c.n)] = j;
                    packed += 1u << (6 * circ);
                }
            }
    nb_cnt[i] = packed;
    if (dropped) {
        atomicOr(err, (uint32_t)PBF_FLAG_FRIENDS_OVERFLOW);
        atomicAdd((unsigned long long*)(err + 2), (unsigned long long)dropped);
    }
}

// compute_scaling.cl:1-113 + pack_data.cl (lambda into x*.w, in place: no
// thread of this kernel reads another particle's w).
__global__ __launch_bounds__(kThreads) void k_lambda(float4* __restrict__ xs, const uint32_t* __restrict__ nb_cnt,
                                                    const uint32_t* __restrict__ nb_idx, float* __restrict__ rho,
                                                    float* __restrict__ lambda, StepConsts c) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= c.n) return;
    const float4 p = xs[i];
    const uint32_t packed = nb_cnt[i];
    const int L = circles_processed<3, 5>(packed);
    float dens = 0.0f, gsum = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
    for (int q = 0; q < L; ++q) {
        const int cq = circle_count(packed, q);
        for (int f = 0; f < cq; ++f) {
            const uint32_t j = nb_idx[friend_slot(q, f, i, c.n)];
            const float4 pj = xs[j];
            const float rx = p.x - pj.x, ry = p.y - pj.y, rz = p.z - pj.z;
            const float r2 = rx * rx + ry * ry + rz * rz;
            if (r2 < c.h2) {
                const float rl = sqrtf(r2);
                const float hr = c.h - rl;
                const float s = c.spiky * hr * hr;
                const float ax = s * rx / rl, ay = s * ry / rl, az = s * rz / rl;
                const float d = c.h2 - r2;
                dens += d * d * d;
                gsum += ax * ax + ay * ay + az * az;
                gx += ax; gy += ay; gz += az;
            }
        }
    }
    dens *= c.poly6;
    rho[i] = dens;
    gsum += gx * gx + gy * gy + gz * gz;
    const float C = (dens / c.rest_density) - 1.0f;
    const float lam = -1.0f * C / (gsum / (c.rest_density * c.rest_density) + c.epsilon);
    lambda[i] = lam;
    reinterpret_cast<float*>(xs + i)[3] = lam;
}

// compute_delta.cl:44-186 + update_predicted.cl; on the last iteration also
// pack_data.cl (density into x*.w) and update_velocities.cl.
template <bool LAST>
__global__ __launch_bounds__(kThreads) void k_delta(const float4* __restrict__ xs, float4* __restrict__ xs_next,
                                                   float4* __restrict__ x, float4* __restrict__ v,
                                                   const float* __restrict__ rho,
                                                   const uint32_t* __restrict__ nb_cnt,
                                                   const uint32_t* __restrict__ nb_idx, StepConsts c) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= c.n) return;
    const float4 p = xs[i];
    const uint32_t packed = nb_cnt[i];
    const int L = circles_processed<1, 2>(packed);
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (int q = 0; q < L; ++q) {
        const int cq = circle_count(packed, q);
        for (int f = 0; f < cq; ++f) {
            const uint32_t j = nb_idx[friend_slot(q, f, i, c.n)];
            const float4 pj = xs[j];
            const float rx = p.x - pj.x, ry = p.y - pj.y, rz = p.z - pj.z;
            const float r2 = rx * rx + ry * ry + rz * rz;
            if (r2 < c.h2) {
                const float rl = sqrtf(r2);
                const float hr = c.h - rl;
                const float ax = rx / rl * hr * hr, ay = ry / rl * hr * hr, az = rz / rl * hr * hr;
                const float d = c.h2 - r2;
                const float p6 = d * d * d;
                const float ratio = p6 / c.poly6_q;
                const float scorr = c.st_k * ratio * ratio * ratio * ratio;
                const float coef = p.w + pj.w + scorr;
                sx += coef * ax; sy += coef * ay; sz += coef * az;
            }
        }
    }
    float fx = p.x + (-c.spiky * sx) / c.rest_density;
    float fy = p.y + (-c.spiky * sy) / c.rest_density;
    float fz = p.z + (-c.spiky * sz) / c.rest_density;
    const float4 prev = x[i];
    if (c.walls) apply_walls(fx, fy, fz, prev, c);
    const float dx = fx - p.x, dy = fy - p.y, dz = fz - p.z;
    float4 o;
    o.x = p.x + dx;
    o.y = p.y + dy;
    o.z = p.z + dz;
    if (!LAST) {
        o.w = p.w;
        xs_next[i] = o;
    } else {
        o.w = rho[i];
        xs_next[i] = o;
        const float vx = (o.x - prev.x) / c.dt, vy = (o.y - prev.y) / c.dt, vz = (o.z - prev.z) / c.dt;
        float* vp = reinterpret_cast<float*>(v + i);
        vp[0] = vx;
        vp[1] = vy;
        vp[2] = vz;
        x[i] = make_float4(o.x, o.y, o.z, sqrtf(vx * vx + vy * vy + vz * vz));
    }
}

// apply_viscosity.cl:1-87, Jacobi: v_j from the snapshot v_in.
__global__ __launch_bounds__(kThreads) void k_viscosity(const float4* __restrict__ xs, const float4* __restrict__ v_in,
                                                       float4* __restrict__ v_out, float4* __restrict__ omega,
                                                       const uint32_t* __restrict__ nb_cnt,
                                                       const uint32_t* __restrict__ nb_idx, StepConsts c) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= c.n) return;
    const float4 p = xs[i];
    const float4 vi = v_in[i];
    const uint32_t packed = nb_cnt[i];
    const int L = circles_processed<1, 2>(packed);
    float vsx = 0.0f, vsy = 0.0f, vsz = 0.0f, ox = 0.0f, oy = 0.0f, oz = 0.0f;
    for (int q = 0; q < L; ++q) {
        const int cq = circle_count(packed, q);
        for (int f = 0; f < cq; ++f) {
            const uint32_t j = nb_idx[friend_slot(q, f, i, c.n)];
            const float4 pj = xs[j];
            const float rx = p.x - pj.x, ry = p.y - pj.y, rz = p.z - pj.z;
            const float r2 = rx * rx + ry * ry + rz * rz;
            if (r2 < c.h2 && fabsf(pj.w) > 1e-8f) {
                const float4 vj = v_in[j];
                const float ux = vj.x - vi.x, uy = vj.y - vi.y, uz = vj.z - vi.z;
                const float d = c.h2 - r2;
                const float rl = sqrtf(r2);
                const float hr = c.h - rl;
                const float ax = rx / rl * hr * hr, ay = ry / rl * hr * hr, az = rz / rl * hr * hr;
                ox += uy * az - uz * ay;
                oy += uz * ax - ux * az;
                oz += ux * ay - uy * ax;
                const float p6 = d * d * d;
                vsx += ux * p6; vsy += uy * p6; vsz += uz * p6;
            }
        }
    }
    float4 vo;
    vo.x = vi.x + c.viscosity * (vsx * c.poly6);
    vo.y = vi.y + c.viscosity * (vsy * c.poly6);
    vo.z = vi.z + c.viscosity * (vsz * c.poly6);
    vo.w = vi.w + c.viscosity * 0.0f;
    v_out[i] = vo;
    omega[i] = make_float4(ox * -c.spiky, oy * -c.spiky, oz * -c.spiky, 0.0f);
}

// apply_vorticity.cl:1-87
__global__ __launch_bounds__(kThreads) void k_vorticity(const float4* __restrict__ xs, float4* __restrict__ v,
                                                       const float4* __restrict__ omega,
                                                       const uint32_t* __restrict__ nb_cnt,
                                                       const uint32_t* __restrict__ nb_idx, StepConsts c) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= c.n) return;
    const float4 p = xs[i];
    const uint32_t packed = nb_cnt[i];
    const int L = circles_processed<1, 2>(packed);
    float ex = 0.0f, ey = 0.0f, ez = 0.0f;
    for (int q = 0; q < L; ++q) {
        const int cq = circle_count(packed, q);
        for (int f = 0; f < cq; ++f) {
            const uint32_t j = nb_idx[friend_slot(q, f, i, c.n)];
            const float4 pj = xs[j];
            const float rx = p.x - pj.x, ry = p.y - pj.y, rz = p.z - pj.z;
            const float r2 = rx * rx + ry * ry + rz * rz;
            if (r2 < c.h2 && fabsf(pj.w) > 1e-8f) {
                const float rl = sqrtf(r2);
                const float hr = c.h - rl;
                const float ax = rx / rl * hr * hr, ay = ry / rl * hr * hr, az = rz / rl * hr * hr;
                const float4 oj = omega[j];
                const float wl = sqrtf(oj.x * oj.x + oj.y * oj.y + oj.z * oj.z);
                ex += wl * ax; ey += wl * ay; ez += wl * az;
            }
        }
    }
    ex *= -c.spiky; ey *= -c.spiky; ez *= -c.spiky;
    const float l = sqrtf(ex * ex + ey * ey + ez * ez);
    if (l > 0.0f) { ex /= l; ey /= l; ez /= l; }
    const float4 oi = omega[i];
    const float cx = ey * oi.z - ez * oi.y;
    const float cy = ez * oi.x - ex * oi.z;
    const float cz = ex * oi.y - ey * oi.x;
    float4 vi = v[i];
    vi.x += (c.vorticity * cx) * c.dt;
    vi.y += (c.vorticity * cy) * c.dt;
    vi.z += (c.vorticity * cz) * c.dt;
    v[i] = vi;
}

inline unsigned blocks_for(uint32_t n, uint32_t per_block) { return (unsigned)((n + per_block - 1) / per_block); }

}  // namespace

hipError_t launch_predict_hash(pbf_ctx* ctx, bool reset_old) {
    track_begin(ctx, "predictPositions");
    hipLaunchKernelGGL(k_predict_hash, dim3(blocks_for(ctx->n, kThreads * kPredictItems)), dim3(kThreads), 0,
                       ctx->stream, ctx->x, ctx->v, ctx->keys_in, ctx->keys_sorted,
                       reinterpret_cast<uint2*>(ctx->cells), reset_old ? 1u : 0u, ctx->sort.hist, ctx->sort.passes,
                       ctx->pause_now, ctx->c, ctx->err);
    track_end(ctx);
    return hipGetLastError();
}

hipError_t launch_predict_only(pbf_ctx* ctx, uint32_t pause) {
    track_begin(ctx, "predictPositions");
    hipLaunchKernelGGL(k_predict_only, dim3(blocks_for(ctx->n, kThreads)), dim3(kThreads), 0, ctx->stream, ctx->x,
                       ctx->v, ctx->xs[0], pause, ctx->c);
    track_end(ctx);
    ctx->xs_cur = 0;
    return hipGetLastError();
}

hipError_t launch_gather(pbf_ctx* ctx) {
    track_begin(ctx, "sortParticles");
    hipLaunchKernelGGL(k_gather, dim3(blocks_for(ctx->n, kThreads)), dim3(kThreads), 0, ctx->stream, ctx->perm,
                       ctx->x, ctx->v, ctx->x_alt, ctx->xs[0], ctx->pause_now, ctx->c, ctx->err);
    track_end(ctx);
    float4* t = ctx->x;
    ctx->x = ctx->x_alt;
    ctx->x_alt = t;
    ctx->xs_cur = 0;
    return hipGetLastError();
}

hipError_t launch_cells(pbf_ctx* ctx) {
    track_begin(ctx, "updateCells");
    hipLaunchKernelGGL(k_cells, dim3(blocks_for(ctx->n, kThreads)), dim3(kThreads), 0, ctx->stream,
                       ctx->keys_sorted, ctx->cells, ctx->n, ctx->c.gbs, ctx->err);
    track_end(ctx);
    return hipGetLastError();
}

hipError_t launch_friends(pbf_ctx* ctx) {
    track_begin(ctx, "buildFriendsList");
    hipLaunchKernelGGL(k_friends, dim3(blocks_for(ctx->n, kThreads)), dim3(kThreads), 0, ctx->stream,
                       ctx->xs[ctx->xs_cur], reinterpret_cast<const uint2*>(ctx->cells), ctx->nb_cnt, ctx->nb_idx,
                       ctx->c, ctx->err);
    track_end(ctx);
    return hipGetLastError();
}

hipError_t launch_lambda(pbf_ctx* ctx, int iter) {
    track_begin(ctx, "computeScaling", iter);
    hipLaunchKernelGGL(k_lambda, dim3(blocks_for(ctx->n, kThreads)), dim3(kThreads), 0, ctx->stream,
                       ctx->xs[ctx->xs_cur], ctx->nb_cnt, ctx->nb_idx, ctx->rho, ctx->lambda, ctx->c);
    track_end(ctx);
    return hipGetLastError();
}

hipError_t launch_delta(pbf_ctx* ctx, int iter, bool last) {
    track_begin(ctx, "computeDelta", iter);
    const int cur = ctx->xs_cur;
    if (last)
        hipLaunchKernelGGL(k_delta<true>, dim3(blocks_for(ctx->n, kThreads)), dim3(kThreads), 0, ctx->stream,
                           ctx->xs[cur], ctx->xs[1 - cur], ctx->x, ctx->v, ctx->rho, ctx->nb_cnt, ctx->nb_idx,
                           ctx->c);
    else
        hipLaunchKernelGGL(k_delta<false>, dim3(blocks_for(ctx->n, kThreads)), dim3(kThreads), 0, ctx->stream,
                           ctx->xs[cur], ctx->xs[1 - cur], ctx->x, ctx->v, ctx->rho, ctx->nb_cnt, ctx->nb_idx,
                           ctx->c);
    track_end(ctx);
    ctx->xs_cur = 1 - cur;
    return hipGetLastError();
}

hipError_t launch_viscosity(pbf_ctx* ctx) {
    track_begin(ctx, "applyViscosity");
    hipLaunchKernelGGL(k_viscosity, dim3(blocks_for(ctx->n, kThreads)), dim3(kThreads), 0, ctx->stream,
                       ctx->xs[ctx->xs_cur], ctx->v, ctx->v_alt, ctx->omega, ctx->nb_cnt, ctx->nb_idx, ctx->c);
    track_end(ctx);
    float4* t = ctx->v;
    ctx->v = ctx->v_alt;
    ctx->v_alt = t;
    return hipGetLastError();
}

hipError_t launch_vorticity(pbf_ctx* ctx) {
    track_begin(ctx, "applyVorticity");
    hipLaunchKernelGGL(k_vorticity, dim3(blocks_for(ctx->n, kThreads)), dim3(kThreads), 0, ctx->stream,
                       ctx->xs[ctx->xs_cur], ctx->v, ctx->omega, ctx->nb_cnt, ctx->nb_idx, ctx->c);
    track_end(ctx);
    return hipGetLastError();
}

}  // namespace pbf
