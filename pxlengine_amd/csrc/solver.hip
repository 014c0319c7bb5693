// Pressure solvers on the matrix of the reference's Gauss-Seidel sweep (fluid.slang:168-182):
// for FLUID cell i, (A p)_i = cnt_i p_i - sum_{FLUID nbrs} p_j (cnt_i = non-SOLID neighbours,
// AIR neighbours are Dirichlet 0), b_i = -div_i h^2 / dt.
//
// Work distribution: every kernel here walks the step's FLUID-cell list (built after
// classification, red cells first then black, each in cell order).  A launch is a fixed grid
// of `cg_blocks` workgroups; workgroup b takes the list slice [n*b/G, n*(b+1)/G), so every
// lane holds liquid and the number of reduction partials is fixed, whatever the liquid's
// shape.
//
// Conjugate gradients with the Jacobi preconditioner M = diag(A), in the Chronopoulos-Gear
// arrangement: both dot products of an iteration come out of ONE fused stencil pass, so an
// iteration is ONE kernel launch:
//
//   scalars:  gamma, delta = sums of the previous pass's workgroup partials;
//             beta = gamma / gamma_old, alpha = gamma / (delta - beta gamma / alpha_old)
//             (first iteration: alpha = gamma / delta, beta = 0)
//   update:   p_i = u_i + beta p_{i-1}      s_i = w_i + beta s_{i-1}
//             x  += alpha p_i               r_{i+1} = r_i - alpha s_i
//             u_{i+1} = M^-1 r_{i+1}        w_{i+1} = A u_{i+1}
//   partials of gamma = (r_{i+1}, u_{i+1}) and delta = (w_{i+1}, u_{i+1}).
//
// Every workgroup sums the previous partials itself, in workgroup order with a fixed tree,
// so all workgroups derive bitwise the same alpha and beta and no finalize launch or
// cross-workgroup synchronisation is needed (the kernel boundary orders the passes).
// Partials and the (alpha, gamma) history ping-pong between two slots by iteration parity.
// w_{i+1} needs u_{i+1} at the six neighbours; the pass rebuilds them from r_i, w_i, s_{i-1}
// read at the neighbours, so r, w and s ping-pong between two buffers each.
//
// Accumulation type Acc = float (plain FP32 CG) or double (the FP64 dot products of the
// mixed-precision solve); vector arithmetic is fp32 in both.
//
// Mixed precision: the FP64 outer loop keeps x64 and r64 = b - A x64 in FP64, solves
// A e = float(r64) from e = 0 with the iterations above, then x64 += e and recomputes r64.
#include "common.hpp"
#include "kernels.hpp"

namespace pxl {

namespace {

constexpr int SBLOCK = 1024;  // lanes per solver workgroup (16 wave64)
constexpr int SWAVES = SBLOCK / WAVE;
// List entries a lane of the MG-CG kernels has in flight at once (their loads are issued
// before any of their stores), within 64 VGPRs (two 1024-lane workgroups per CU).  Measured on
// the 64-request batch: two entries make spmv 15% faster (four: 60% slower).  With the global
// (not flat) loads of the batched kernels, two entries make update 5-10% faster on the
// 512-request batch and leave it unchanged on the 64-request batch (four: 20% slower there).
constexpr int SPMV_MLP = 2, UPD_MLP = 2;

// Fields of a list entry's stencil word (see DevSim::fl_word).
__device__ __forceinline__ bool w_fluid(uint32_t w, int n) { return (w >> n) & 1u; }
__device__ __forceinline__ float w_diag(uint32_t w) { return (float)((w >> 6) & 7u); }
__device__ __forceinline__ float w_ndiag(uint32_t w, int n) { return (float)((w >> (9 + 3 * n)) & 7u); }

// Jacobi preconditioner u = r * (1 / diag) with the hardware reciprocal (a preconditioner
// need not be exact; 0 where the diagonal is 0).  The diagonal is a small integer 1..6.
__device__ __forceinline__ float precond(float r, float dg) {
    return dg > 0.0f ? r * __builtin_amdgcn_rcpf(dg) : 0.0f;
}

__device__ __forceinline__ void nb_offsets(const DevSim &S, int32_t off[6]) {
    const int32_t sy = S.nx, sz = S.nx * S.ny;
    off[0] = -1; off[1] = 1; off[2] = -sy; off[3] = sy; off[4] = -sz; off[5] = sz;
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup i runs on XCD i % 8), and each
// XCD has its own L2.  Slice s of the list goes to workgroup xcd_slot(s), so each XCD gets
// one contiguous eighth of the list and the neighbour rows and planes a slice reads are
// mostly in the same L2.
constexpr uint32_t NXCD = 8;

// Workgroup b's (of nb) slice of list entries [lo, hi).
__device__ __forceinline__ void list_slice(uint32_t lo, uint32_t hi, uint32_t b_in, uint32_t nb, uint32_t &beg,
                                           uint32_t &end) {
    const uint64_t n = hi - lo, g = nb;
    uint64_t b = b_in;
    if (g % NXCD == 0) b = (b % NXCD) * (g / NXCD) + b / NXCD;
    beg = lo + (uint32_t)(n * b / g);
    end = lo + (uint32_t)(n * (b + 1) / g);
}

template <typename Acc>
__device__ __forceinline__ Acc *partials(const DevSim &S);
template <>
__device__ __forceinline__ float *partials<float>(const DevSim &S) { return S.part32; }
template <>
__device__ __forceinline__ double *partials<double>(const DevSim &S) { return S.part64; }

// Stores the sums of NV values of workgroup b (of nb) at part[k * nb + b] (plain stores: they
// are consumed by the next kernel on the stream).
template <typename T, int NV>
__device__ __forceinline__ void store_partials(const T (&v)[NV], T *part, uint32_t b, uint32_t nb) {
    __shared__ T sh[SWAVES * NV];
#pragma unroll
    for (int k = 0; k < NV; k++) {
        T s = block_sum<T, SWAVES>(v[k], sh + SWAVES * k);
        if (threadIdx.x == 0) part[(size_t)k * nb + b] = s;
    }
}

// Sum of the nb workgroup partials of value k, in workgroup order with a fixed tree.
template <typename T>
__device__ __forceinline__ T sum_partials(const T *part, uint32_t nb, int k, T *sh) {
    T a = 0;
    for (uint32_t i = threadIdx.x; i < nb; i += BLOCK) a += part[(size_t)k * nb + i];
    return block_sum(a, sh);
}

// The same total computed by each wave on its own (lane l adds entries l, l+64, ...; then a
// butterfly), so no workgroup barrier is needed; every wave of every workgroup gets the
// same bits.
template <typename T>
__device__ __forceinline__ T wave_sum_partials(const T *part, uint32_t nb, int k) {
    T a = 0;
    for (uint32_t i = threadIdx.x & (WAVE - 1); i < nb; i += WAVE) a += part[(size_t)k * nb + i];
    return wave_sum(a);
}

// ---- init 1: r0 = b - A x (pressure system, x = warm-started pressure) or, for the inner
// solve of the refinement, r0 = float(r64) (its iterate starts at 0: the MG-CG's first spmv
// clears it).
__device__ __forceinline__ void body_cg_residual(const DevSim &S, const uint32_t *flc, float *__restrict__ x,
                                                 float *__restrict__ r, int inner, uint32_t b, uint32_t nb) {
    uint32_t beg, end;
    list_slice(0, flc[1], b, nb, beg, end);
    int32_t off[6];
    nb_offsets(S, off);
    for (uint32_t t = beg + threadIdx.x; t < end; t += SBLOCK) {
        if (inner) {  // the inner solve's r: colour-split (DevSim::fls_list) and in list order
            const float rv = (float)S.r64[S.fl_list[t]];
            r[S.fls_list[t]] = rv;
            S.cg_rt[t] = rv;
            continue;
        }
        const uint32_t id = S.fn_list[t];
        const uint32_t wd = S.fn_word[t];
        float sum = 0.0f;
#pragma unroll
        for (int n = 0; n < 6; n++)
            if (w_fluid(wd, n)) sum = sum + x[id + off[n]];
        const float ax = w_diag(wd) * x[id] - sum;
        const float bb = -(S.div[id] * S.h * S.h / S.dt);
        r[id] = bb - ax;
    }
}
__global__ __launch_bounds__(SBLOCK) void k_cg_residual(DevSim S, float *__restrict__ x, float *__restrict__ r,
                                                       int inner) {
    body_cg_residual(S, S.fl_count, x, r, inner, blockIdx.x, gridDim.x);
}

// ---- init 2: u0 = M^-1 r0, w0 = A u0; p_{-1} = s_{-1} = 0; partials gamma0, delta0 (slot 0).
template <typename Acc>
__global__ __launch_bounds__(SBLOCK) void k_cg_start(DevSim S, const float *__restrict__ r, float *__restrict__ w,
                                                    float *__restrict__ s_prev) {
    uint32_t beg, end;
    list_slice(0, S.fl_count[1], blockIdx.x, gridDim.x, beg, end);
    int32_t off[6];
    nb_offsets(S, off);
    Acc g = 0, d = 0;
    for (uint32_t t = beg + threadIdx.x; t < end; t += SBLOCK) {
        const uint32_t id = S.fn_list[t];
        const uint32_t wd = S.fn_word[t];
        const float uc = precond(r[id], w_diag(wd));
        float sum = 0.0f;
#pragma unroll
        for (int n = 0; n < 6; n++)
            if (w_fluid(wd, n)) sum = sum + precond(r[id + off[n]], w_ndiag(wd, n));
        const float wc = w_diag(wd) * uc - sum;
        w[id] = wc;
        S.cg_p[id] = 0.0f;
        s_prev[id] = 0.0f;
        g += (Acc)r[id] * (Acc)uc;
        d += (Acc)wc * (Acc)uc;
    }
    const Acc v[2] = {g, d};
    store_partials<Acc, 2>(v, partials<Acc>(S), blockIdx.x, gridDim.x);
}

// Scalars of iteration `it` from the previous pass's partials (identical in every workgroup).
template <typename Acc>
__device__ __forceinline__ void cg_scalars(const DevSim &S, uint32_t it, uint32_t b_self, uint32_t nb, float &alpha,
                                           float &beta) {
    const Acc *part = partials<Acc>(S) + (size_t)(it & 1) * 2 * nb;
    const Acc g = wave_sum_partials<Acc>(part, nb, 0);
    const Acc d = wave_sum_partials<Acc>(part, nb, 1);
    Acc a = 0, b = 0;
    if (it == 0) {
        a = (g != (Acc)0 && d != (Acc)0) ? g / d : (Acc)0;
    } else {
        const double *prev = S.scal + SC_COUNT * (it & 1);
        const Acc g_old = (Acc)prev[SC_RZ], a_old = (Acc)prev[SC_ALPHA];
        if (g != (Acc)0 && g_old != (Acc)0 && a_old != (Acc)0) {
            b = g / g_old;
            const Acc den = d - b * g / a_old;
            a = den != (Acc)0 ? g / den : (Acc)0;
        }
    }
    if (b_self == 0 && threadIdx.x == 0) {
        double *next = S.scal + SC_COUNT * ((it + 1) & 1);
        next[SC_ALPHA] = (double)a;
        next[SC_BETA] = (double)b;
        next[SC_RZ] = (double)g;
    }
    alpha = (float)a;
    beta = (float)b;
}

// ---- one fused iteration (see the file comment).  All neighbour loads are issued
// unconditionally (a neighbour that is not FLUID reads the cell itself; FLUID cells are never
// on the wall layer, so every neighbour index is in bounds) and selected afterwards, so a lane
// keeps all its loads in flight.
template <typename Acc>
__global__ __launch_bounds__(SBLOCK) void k_cg_iter(DevSim S, uint32_t it, float *__restrict__ x,
                                                   const float *__restrict__ r_in, float *__restrict__ r_out,
                                                   const float *__restrict__ w_in, float *__restrict__ w_out,
                                                   const float *__restrict__ s_prev, float *__restrict__ s_cur) {
    uint32_t beg, end;
    list_slice(0, S.fl_count[1], blockIdx.x, gridDim.x, beg, end);
    const float *__restrict__ pv = S.cg_p;
    int32_t off[6];
    nb_offsets(S, off);
    float alpha, beta;
    cg_scalars<Acc>(S, it, blockIdx.x, gridDim.x, alpha, beta);
    Acc g = 0, d = 0;
    for (uint32_t t = beg + threadIdx.x; t < end; t += SBLOCK) {
        const uint32_t c = S.fn_list[t];
        const uint32_t wd = S.fn_word[t];
        uint32_t q[6];
#pragma unroll
        for (int n = 0; n < 6; n++) q[n] = w_fluid(wd, n) ? c + off[n] : c;
        const float ri = r_in[c], wi = w_in[c], si = s_prev[c], pi = pv[c], xi = x[c];
        float rq[6], wq[6], sq[6];
#pragma unroll
        for (int n = 0; n < 6; n++) {
            rq[n] = r_in[q[n]];
            wq[n] = w_in[q[n]];
            sq[n] = s_prev[q[n]];
        }
        const float dg = w_diag(wd);
        const float pc = precond(ri, dg) + beta * pi;
        const float sc = wi + beta * si;
        const float rn = ri - alpha * sc;
        const float un = precond(rn, dg);
        float sum = 0.0f;
#pragma unroll
        for (int n = 0; n < 6; n++) {
            const float un_n = precond(rq[n] - alpha * (wq[n] + beta * sq[n]), w_ndiag(wd, n));
            sum = w_fluid(wd, n) ? sum + un_n : sum;
        }
        const float wn = dg * un - sum;
        x[c] = xi + alpha * pc;
        S.cg_p[c] = pc;
        s_cur[c] = sc;
        r_out[c] = rn;
        w_out[c] = wn;
        g += (Acc)rn * (Acc)un;
        d += (Acc)wn * (Acc)un;
    }
    const Acc v[2] = {g, d};
    store_partials<Acc, 2>(v, partials<Acc>(S) + (size_t)((it + 1) & 1) * 2 * gridDim.x, blockIdx.x, gridDim.x);
}

// ---- multigrid-preconditioned CG (Chronopoulos-Gear with u = M r from the V-cycle).
// r = cg_r0 and u = cg_r1, which the V-cycle reads with neighbours, are colour-split (an entry's
// index is fls_list[t], its neighbours' split_offsets away); r is kept in list order as well
// (cg_rt) for the kernels that read it only at the entry itself.  w = cg_w0, s = cg_s0, p = cg_p and
// the iterate x = cg_x are only ever read at the entry itself, so they are stored in list
// order (index t): runs of a short row or a small grid then leave no gaps in their cache lines
// (the colour-split layout touches whole lines even for a few FLUID cells of a row).  spmv and
// update walk the colour-split list (red cells, then black, each in cell order); the last
// update of a pass also stores x colour-split in e, for the FP64 refinement.
// spmv: w = A u and the partials of gamma = (r, u), delta = (w, u) into slot `slot`
// (first: also p_{-1} = s_{-1} = 0 and x = 0).
template <typename Acc>
__device__ __forceinline__ void body_mgcg_spmv(const DevSim &S, const uint32_t *flc, uint32_t slot, int first,
                                               uint32_t b, uint32_t nb) {
    uint32_t beg, end;
    list_slice(0, flc[1], b, nb, beg, end);
    const float *__restrict__ u = S.cg_r1;
    Acc g = 0, d = 0;
    // a lane takes entries t0 + k SBLOCK (k < SPMV_MLP) with the loads of all of them in flight
    // before any store; it still meets its entries in increasing order, so the partial sums
    // add in the same order as one entry at a time
    for (uint32_t t0 = beg + threadIdx.x; t0 < end; t0 += SPMV_MLP * SBLOCK) {
        uint32_t c[SPMV_MLP], wd[SPMV_MLP];
#pragma unroll
        for (int k = 0; k < SPMV_MLP; k++) {
            const uint32_t t = t0 + k * SBLOCK;
            c[k] = S.fls_list[t < end ? t : t0];
            wd[k] = S.fl_word[t < end ? t : t0];  // loaded unconditionally, then dropped
            wd[k] = t < end ? wd[k] : 0u;  // past the slice: no neighbours, no store
        }
        float un[SPMV_MLP][6], uc[SPMV_MLP], rc[SPMV_MLP];
#pragma unroll
        for (int k = 0; k < SPMV_MLP; k++) {
            int32_t off[6];
            split_offsets(S, wd[k], off);
#pragma unroll
            for (int n = 0; n < 6; n++) {
                const bool f = w_fluid(wd[k], n);
                const float v = u[f ? c[k] + off[n] : c[k]];
                un[k][n] = f ? v : 0.0f;
            }
            uc[k] = u[c[k]];
            rc[k] = S.cg_rt[t0 + k * SBLOCK < end ? t0 + k * SBLOCK : t0];
        }
#pragma unroll
        for (int k = 0; k < SPMV_MLP; k++) {
            if (t0 + k * SBLOCK >= end) break;
            float sum = 0.0f;
#pragma unroll
            for (int n = 0; n < 6; n++)
                if (w_fluid(wd[k], n)) sum = sum + un[k][n];
            const float wc = w_diag(wd[k]) * uc[k] - sum;
            const uint32_t t = t0 + k * SBLOCK;
            S.cg_w0[t] = wc;
            if (first) {
                S.cg_p[t] = 0.0f;
                S.cg_s0[t] = 0.0f;
                S.cg_x[t] = 0.0f;
            }
            g += (Acc)rc[k] * (Acc)uc[k];
            d += (Acc)wc * (Acc)uc[k];
        }
    }
    const Acc v[2] = {g, d};
    store_partials<Acc, 2>(v, partials<Acc>(S) + (size_t)slot * 2 * nb, b, nb);
}

// update of iteration it: scalars from slot it & 1, then p, s, x, r (no neighbours); `last`:
// x goes to e (colour-split) instead of cg_x.
template <typename Acc>
__device__ __forceinline__ void body_mgcg_update(const DevSim &S, const uint32_t *flc, uint32_t it, int last,
                                                 float *__restrict__ e, uint32_t b, uint32_t nb) {
    uint32_t beg, end;
    list_slice(0, flc[1], b, nb, beg, end);
    float alpha, beta;
    cg_scalars<Acc>(S, it, b, nb, alpha, beta);
    for (uint32_t t0 = beg + threadIdx.x; t0 < end; t0 += UPD_MLP * SBLOCK) {
        uint32_t c[UPD_MLP];
#pragma unroll
        for (int k = 0; k < UPD_MLP; k++) {
            const uint32_t t = t0 + k * SBLOCK;
            c[k] = S.fls_list[t < end ? t : t0];
        }
        float uu[UPD_MLP], pp[UPD_MLP], ww[UPD_MLP], ss[UPD_MLP], xx[UPD_MLP], rr[UPD_MLP];
#pragma unroll
        for (int k = 0; k < UPD_MLP; k++) {
            const uint32_t t = min(t0 + k * SBLOCK, end - 1u);
            uu[k] = S.cg_r1[c[k]];
            pp[k] = S.cg_p[t];
            ww[k] = S.cg_w0[t];
            ss[k] = S.cg_s0[t];
            xx[k] = S.cg_x[t];
            rr[k] = S.cg_rt[t];
        }
#pragma unroll
        for (int k = 0; k < UPD_MLP; k++) {
            const uint32_t t = t0 + k * SBLOCK;
            if (t >= end) break;
            const float pc = uu[k] + beta * pp[k];
            const float sc = ww[k] + beta * ss[k];
            S.cg_p[t] = pc;
            S.cg_s0[t] = sc;
            const float xn = xx[k] + alpha * pc;
            if (last) e[c[k]] = xn;
            else S.cg_x[t] = xn;
            const float rn = rr[k] - alpha * sc;
            S.cg_r0[c[k]] = rn;
            S.cg_rt[t] = rn;
        }
    }
}
template <typename Acc>
__global__ __launch_bounds__(SBLOCK) void k_mgcg_spmv(DevSim S, uint32_t slot, int first) {
    body_mgcg_spmv<Acc>(S, S.fl_count, slot, first, blockIdx.x, gridDim.x);
}
template <typename Acc>
__global__ __launch_bounds__(SBLOCK) void k_mgcg_update(DevSim S, uint32_t it, int last, float *__restrict__ e) {
    body_mgcg_update<Acc>(S, S.fl_count, it, last, e, blockIdx.x, gridDim.x);
}

// ---- FP64 outer loop.  The solution is double-buffered (x_src -> x_dst) so the update pass
// can rebuild neighbours' x_src + e from values no lane has overwritten.
__device__ __forceinline__ double b64_of(const DevSim &S, uint32_t id) {
    return -((double)S.div[id] * (double)S.h * (double)S.h / (double)S.dt);
}

// init: x64 = pressure, r64 = b - A x64; partials of rr and bb.
__device__ __forceinline__ void body_mixed_init(const DevSim &S, const uint32_t *flc, double *__restrict__ x64,
                                                uint32_t b_, uint32_t nb) {
    uint32_t beg, end;
    list_slice(0, flc[1], b_, nb, beg, end);
    int32_t off[6];
    nb_offsets(S, off);
    double rr = 0, bb = 0;
    // FLUID neighbours are read with masked raw buffer loads and added by select (the sum is the
    // same as adding only those), so a lane's loads are in flight together
    const Buf bp = buf_of(S.pressure, 4u * S.ncells);
    for (uint32_t t = beg + threadIdx.x; t < end; t += SBLOCK) {
        const uint32_t id = S.fn_list[t];
        const uint32_t wd = S.fn_word[t];
        float pn[6];
#pragma unroll
        for (int n = 0; n < 6; n++) pn[n] = buf_f32(bp, id + off[n], w_fluid(wd, n));
        double sum = 0.0;
#pragma unroll
        for (int n = 0; n < 6; n++) sum = w_fluid(wd, n) ? sum + (double)pn[n] : sum;
        const double xc = (double)S.pressure[id];
        const double b = b64_of(S, id);
        const double r = b - ((double)w_diag(wd) * xc - sum);
        x64[id] = xc;
        S.r64[id] = r;
        rr += r * r;
        bb += b * b;
    }
    const double v[2] = {rr, bb};
    store_partials<double, 2>(v, S.part64, b_, nb);
}
__global__ __launch_bounds__(SBLOCK) void k_mixed_init(DevSim S, double *__restrict__ x64) {
    body_mixed_init(S, S.fl_count, x64, blockIdx.x, gridDim.x);
}

// update: x_dst = x_src + e (FP64) at the cell and its neighbours, r64 = b - A x_dst; rr.
// The pressure field gets float(x_dst), so it holds the latest iterate after every pass.
__device__ __forceinline__ void body_mixed_update(const DevSim &S, const uint32_t *flc,
                                                  const double *__restrict__ xsrc, double *__restrict__ xdst,
                                                  const float *__restrict__ e, uint32_t b, uint32_t nb) {
    uint32_t beg, end;
    list_slice(0, flc[1], b, nb, beg, end);
    int32_t off[6];
    nb_offsets(S, off);
    double rr = 0;
    // FLUID neighbours are read with masked raw buffer loads and added by select (the sum is the
    // same as adding only those), so a lane's loads are in flight together
    const Buf bx = buf_of(xsrc, 8u * S.ncells), be = buf_of(e, 8u * S.hsplit);
    for (uint32_t t = beg + threadIdx.x; t < end; t += SBLOCK) {
        const uint32_t id = S.fn_list[t], sid = S.fns_list[t];
        const uint32_t wd = S.fn_word[t];
        int32_t soff[6];
        split_offsets(S, wd, soff);  // e is the inner solve's colour-split solution
        double xn[6];
        float en[6];
#pragma unroll
        for (int n = 0; n < 6; n++) {
            xn[n] = buf_f64(bx, id + off[n], w_fluid(wd, n));
            en[n] = buf_f32(be, sid + soff[n], w_fluid(wd, n));
        }
        double sum = 0.0;
#pragma unroll
        for (int n = 0; n < 6; n++) sum = w_fluid(wd, n) ? sum + (xn[n] + (double)en[n]) : sum;
        const double xc = xsrc[id] + (double)e[sid];
        const double r = b64_of(S, id) - ((double)w_diag(wd) * xc - sum);
        xdst[id] = xc;
        S.pressure[id] = (float)xc;
        S.r64[id] = r;
        rr += r * r;
    }
    const double v[1] = {rr};
    store_partials<double, 1>(v, S.part64, b, nb);
}
__global__ __launch_bounds__(SBLOCK) void k_mixed_update(DevSim S, const double *__restrict__ xsrc,
                                                        double *__restrict__ xdst, const float *__restrict__ e) {
    body_mixed_update(S, S.fl_count, xsrc, xdst, e, blockIdx.x, gridDim.x);
}

// Totals of the refinement norms into norms[0] = rr (and norms[1] = bb when nv = 2).
__device__ __forceinline__ void body_norms_finalize(const DevSim &S, uint32_t nb, int nv) {
    __shared__ double sh[8];
    const double rr = sum_partials<double>(S.part64, nb, 0, sh);
    double bb = 0;
    if (nv > 1) bb = sum_partials<double>(S.part64, nb, 1, sh + 4);
    if (threadIdx.x != 0) return;
    S.norms[0] = rr;
    if (nv > 1) S.norms[1] = bb;
}
__global__ __launch_bounds__(BLOCK) void k_norms_finalize(DevSim S, uint32_t nb, int nv) {
    body_norms_finalize(S, nb, nv);
}

// ---- batched (the solving simulations of a batch; x = e, r = cg_r0; the fluid-list counts
// come from the batch's count slots)
__global__ __launch_bounds__(SBLOCK) void kb_cg_residual(WorkList W) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, const uint32_t *cnt) {
        body_cg_residual(S, cnt, S.e, S.cg_r0, 1, b, S.cg_blocks);
    });
}
__global__ __launch_bounds__(SBLOCK) void kb_mgcg_spmv(WorkList W, uint32_t slot, int first) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, const uint32_t *cnt) {
        body_mgcg_spmv<double>(S, cnt, slot, first, b, S.cg_blocks);
    });
}
__global__ __launch_bounds__(SBLOCK) void kb_mgcg_update(WorkList W, uint32_t it, int last) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, const uint32_t *cnt) {
        body_mgcg_update<double>(S, cnt, it, last, S.e, b, S.cg_blocks);
    });
}
__global__ __launch_bounds__(SBLOCK) void kb_mixed_init(WorkList W) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, const uint32_t *cnt) {
        body_mixed_init(S, cnt, S.x64a, b, S.cg_blocks);
    });
}
__global__ __launch_bounds__(SBLOCK) void kb_mixed_update(WorkList W, int cur) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, const uint32_t *cnt) {
        body_mixed_update(S, cnt, cur ? S.x64b : S.x64a, cur ? S.x64a : S.x64b, S.e, b, S.cg_blocks);
    });
}
__global__ __launch_bounds__(BLOCK) void kb_norms_finalize(WorkList W, int nv) {
    for_each_work(W, [&](const DevSim &S, uint32_t, const uint32_t *) { body_norms_finalize(S, (uint32_t)S.cg_blocks, nv); });
}

// ---- pressure_rb (fluid.slang:168-182): red-black half sweeps over the FLUID cells, up to
// RB_HALO of them per launch (temporal blocking).  The update of a FLUID cell is the
// reference's, in its fp32 order: p = (sum - div h h / dt) / cnt, with sum = 0 + the FLUID
// neighbours' pressures in the order -x, +x, -y, +y, -z, +z and cnt its non-SOLID neighbours;
// a cell with cnt = 0 keeps its value.
//
// A workgroup owns a CX x CY x CZ core of the grid and holds the core plus RB_HALO cells on
// every side (its region) in LDS, the pressures split by colour.  Half sweep s (1-based) of a
// launch updates the region's FLUID cells of its colour at least s cells from the region's
// faces; their six neighbours are then at least s - 1 from the faces and hold the values
// after half sweep s - 1, so after s <= RB_HALO half sweeps the core holds what s whole-grid
// half sweeps give.  The sum runs over all six neighbours: cells that are not FLUID hold
// pressure 0 (the cell pass zeroes them, and no sweep writes them), and a sum that starts at
// +0 is never -0, so adding a zero leaves it unchanged and 0 + p0 + ... + p5 equals the sum
// over the FLUID neighbours bit for bit.
//
// The FLUID cells do not change during the solve, so one launch per solve lists each core's
// region cells to update (FLUID, at least one cell inside the region, with a non-SOLID
// neighbour) per colour, ordered by distance from the faces (4 or more, 3, 2, 1): half sweep
// s then walks the prefix at distance >= s.  A sweep launch reads `pin` and writes its core
// to `pout`, so no workgroup reads a value another has already moved past.  A core with no
// listed cell keeps its values; it is copied by the first launch (whose `pout` holds values
// of an earlier solve) and skipped by the others.
constexpr int RB_HALO = 4;

template <int CX, int CY, int CZ, int NT>
struct RbTile {
    static constexpr int RX = CX + 2 * RB_HALO, RY = CY + 2 * RB_HALO, RZ = CZ + 2 * RB_HALO;
    static constexpr int HX = RX / 2;                 // cells of one colour in a region row
    static constexpr int NC = HX * RY * RZ;           // cells of one colour in the region
    static constexpr int SL = (NC + NT - 1) / NT;     // of them per lane
    static constexpr int NQ = RX * RY * RZ;           // region cells
    static constexpr int LMAX = ((RX - 2) * (RY - 2) * (RZ - 2) + 1) / 2;  // listed cells of a colour, at most
    static constexpr int K = (LMAX + NT - 1) / NT;    // of them per lane
    static_assert(CX % TILE == 0 && CY % TILE == 0 && CZ % TILE == 0, "cores are whole 8^3 tiles");
    static_assert(RX % 2 == 0, "region rows split evenly by colour");
    static_assert(NC < (1 << 13), "13-bit region indices");
};
// The core shape of reference mode: 8 x 8 x 16 cells.  Sweep launches run 1024 lanes per core
// when the grid has at most RB_WIDE_CORES cores (64^3: 256 cores, one per CU) and 512 above
// (128^3: 2048 cores; measured 14% faster there, 9% slower at 64^3).
constexpr int RB_CX = 8, RB_CY = 8, RB_CZ = 16, RB_NT = 1024;
constexpr uint32_t RB_WIDE_CORES = 1024;
using RbCore = RbTile<RB_CX, RB_CY, RB_CZ, RB_NT>;
static_assert(RbTile<RB_CX, RB_CY, RB_CZ, 512>::LMAX == RbCore::LMAX, "one list layout for both widths");
// list entry: bits 0-12 the region index of its colour, bit 13 x & 1, bits 14-16 the diagonal

__device__ __forceinline__ void rb_core_of(const DevSim &S, uint32_t core, int cx, int cy, int cz, int &x0, int &y0,
                                           int &z0) {
    const uint32_t ncx = ((uint32_t)S.nx + cx - 1) / cx, ncy = ((uint32_t)S.ny + cy - 1) / cy;
    x0 = (int)(core % ncx) * cx;
    y0 = (int)((core / ncx) % ncy) * cy;
    z0 = (int)(core / (ncx * ncy)) * cz;
}

template <int CX, int CY, int CZ, int NT>
__global__ __launch_bounds__(NT) void k_rbgs_lists(DevSim S) {
    using T = RbTile<CX, CY, CZ, NT>;
    constexpr int H = RB_HALO, HX = T::HX, RX = T::RX, RY = T::RY, RZ = T::RZ, NC = T::NC, SL = T::SL, LMAX = T::LMAX;
    __shared__ uint32_t cnt[2][4], from[2][4];
    const uint32_t core = blockIdx.x;
    int x0, y0, z0;
    rb_core_of(S, core, CX, CY, CZ, x0, y0, z0);
    const int tid = (int)threadIdx.x, lane = tid & (WAVE - 1);
    if (tid < 8) cnt[tid >> 2][tid & 3] = 0u;
    const int ox = x0 - H, oy = y0 - H, oz = z0 - H;
    const uint32_t ncell = (uint32_t)S.ncells;
    const Buf bd = buf_of(S.div, 4u * ncell);
    const __amdgpu_buffer_rsrc_t bs = __builtin_amdgcn_make_buffer_rsrc(S.sten, (short)0, (int)(2u * ncell), 0x00020000);
    uint32_t st[2][SL];
    float dv[2][SL];
    int lvl[2][SL];
#pragma unroll
    for (int m = 0; m < SL; m++) {
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const int u = tid + NT * m, row = u / HX, y = row % RY, z = row / RY, x = 2 * (u % HX) + ((c ^ y ^ z) & 1);
            const int gx = ox + x, gy = oy + y, gz = oz + z;
            const bool in = u < NC && gx >= 0 && gy >= 0 && gz >= 0 && gx < S.nx && gy < S.ny && gz < S.nz;
            const uint32_t id = in ? (uint32_t)gx + (uint32_t)S.nx * ((uint32_t)gy + (uint32_t)S.ny * (uint32_t)gz) : 0u;
            st[c][m] = __builtin_amdgcn_raw_buffer_load_b16(bs, (int)(in ? 2u * id : 2u * ncell), 0, 0);
            dv[c][m] = buf_f32(bd, id, in);
            lvl[c][m] = in ? min(min(min(x, RX - 1 - x), min(y, RY - 1 - y)), min(z, RZ - 1 - z)) : 0;
        }
    }
    __syncthreads();
    // ranks within (colour, distance) classes: one LDS atomic per wave and class
    uint32_t rank[2][SL];
#pragma unroll
    for (int m = 0; m < SL; m++) {
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const uint32_t s = st[c][m];
            const bool listed = (s & STEN_FLUID) && lvl[c][m] >= 1 && ((s >> 8) & 7u) != 0u;
            const int cls = 4 - min(lvl[c][m], 4);
            rank[c][m] = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint64_t mask = __ballot(listed && cls == k);
                if (mask == 0) continue;
                const int leader = __ffsll((long long)mask) - 1;
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(&cnt[c][k], (uint32_t)__popcll(mask));
                base = __shfl(base, leader, WAVE);
                if (listed && cls == k) rank[c][m] = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
            }
        }
    }
    __syncthreads();
    if (tid < 2) {
        uint32_t a = 0;
        for (int k = 0; k < 4; k++) {
            from[tid][k] = a;
            a += cnt[tid][k];
        }
    }
    __syncthreads();
    uint32_t *lw = S.rb_list + (size_t)core * 2 * LMAX;
    float *lb = S.rb_b + (size_t)core * 2 * LMAX;
    const float h = S.h, dt = S.dt;
#pragma unroll
    for (int m = 0; m < SL; m++) {
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const uint32_t s = st[c][m];
            const bool listed = (s & STEN_FLUID) && lvl[c][m] >= 1 && ((s >> 8) & 7u) != 0u;
            if (listed) {
                const int u = tid + NT * m, row = u / HX, y = row % RY, z = row / RY;
                const uint32_t x1 = (uint32_t)((c ^ y ^ z) & 1);
                const uint32_t pos = c * LMAX + from[c][4 - min(lvl[c][m], 4)] + rank[c][m];
                lw[pos] = (uint32_t)u | (x1 << 13) | (((s >> 8) & 7u) << 14);
                lb[pos] = dv[c][m] * h * h / dt;
            }
        }
    }
    if (tid < 8) S.rb_count[core * 8 + tid] = cnt[tid >> 2][tid & 3];
}

template <int CX, int CY, int CZ, int NT>
__global__ __launch_bounds__(NT) void k_rbgs_tile(DevSim S, const float *pin, float *pout, uint32_t first,
                                                  uint32_t nsweep, int copy_idle) {
    using T = RbTile<CX, CY, CZ, NT>;
    constexpr int H = RB_HALO, HX = T::HX, RX = T::RX, RY = T::RY, NC = T::NC, NQ = T::NQ, LMAX = T::LMAX, K = T::K;
    __shared__ float P[2 * NC];
    uint32_t core = blockIdx.x;
    const uint32_t g = gridDim.x;
    if (g % NXCD == 0) core = (core % NXCD) * (g / NXCD) + core / NXCD;  // neighbouring cores on one XCD
    int x0, y0, z0;
    rb_core_of(S, core, CX, CY, CZ, x0, y0, z0);
    const int tid = (int)threadIdx.x;
    const int ox = x0 - H, oy = y0 - H, oz = z0 - H;
    const uint32_t opar = (uint32_t)(ox + oy + oz) & 1u;
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    // list lengths by class (distance 4+, 3, 2, 1) and colour
    uint32_t cn[2][4];
#pragma unroll
    for (int k = 0; k < 8; k++) cn[k >> 2][k & 3] = S.rb_count[core * 8 + k];
    // the region's pressures (x fastest: whole rows of the grid)
    const Buf bp = buf_of(pin, 4u * (uint32_t)S.ncells);
    constexpr int QS = (NQ + NT - 1) / NT;
    float pv[QS];
#pragma unroll
    for (int m = 0; m < QS; m++) {
        const int q = tid + NT * m, x = q % RX, y = (q / RX) % RY, z = q / (RX * RY);
        const int gx = ox + x, gy = oy + y, gz = oz + z;
        const bool in = q < NQ && gx >= 0 && gy >= 0 && gz >= 0 && gx < nx && gy < ny && gz < nz;
        pv[m] = buf_f32(bp, in ? (uint32_t)gx + (uint32_t)nx * ((uint32_t)gy + (uint32_t)ny * (uint32_t)gz) : 0u, in);
    }
    const uint32_t ntot[2] = {cn[0][0] + cn[0][1] + cn[0][2] + cn[0][3], cn[1][0] + cn[1][1] + cn[1][2] + cn[1][3]};
    const bool idle = cn[0][0] + cn[1][0] == 0u;  // no listed core cell
    if (idle && !copy_idle) return;
    // this lane's list entries
    const uint32_t *lw = S.rb_list + (size_t)core * 2 * LMAX;
    const float *lb = S.rb_b + (size_t)core * 2 * LMAX;
    constexpr uint32_t SAFE = (uint32_t)(HX * RY + 1);  // lanes past the list read around an interior cell
    uint32_t w[2][K];
    float bt[2][K];
#pragma unroll
    for (int c = 0; c < 2; c++) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint32_t t = (uint32_t)(tid + NT * k);
            const bool ok = t < ntot[c];
            w[c][k] = ok ? ld_global(lw, c * LMAX + t) : SAFE;
            bt[c][k] = ok ? ld_global(lb, c * LMAX + t) : 0.0f;
        }
    }
#pragma unroll
    for (int m = 0; m < QS; m++) {
        const int q = tid + NT * m, x = q % RX, y = (q / RX) % RY, z = q / (RX * RY);
        if (q < NQ) P[((x + y + z) & 1) * NC + (y + RY * z) * HX + (x >> 1)] = pv[m];
    }
    __syncthreads();

    const int wave0 = tid & ~(WAVE - 1);
    for (uint32_t j = 0; j < nsweep; j++) {
        const uint32_t lc = ((first + j) & 1u) ^ opar;
        uint32_t n = 0;  // entries at distance > j
#pragma unroll
        for (int k = 0; k < 4; k++) n += (k < 4 - (int)j) ? (lc ? cn[1][k] : cn[0][k]) : 0u;
        const float *Q = P + (lc ^ 1u) * NC;
        float *Pc = P + lc * NC;
        float nb[K][6];
#pragma unroll
        for (int k = 0; k < K; k++) {
            if ((uint32_t)(wave0 + NT * k) < n) {
                const uint32_t e = lc ? w[1][k] : w[0][k];
                const int u = (int)(e & 0x1fffu), x1 = (int)((e >> 13) & 1u);
                nb[k][0] = Q[u - 1 + x1];
                nb[k][1] = Q[u + x1];
                nb[k][2] = Q[u - HX];
                nb[k][3] = Q[u + HX];
                nb[k][4] = Q[u - HX * RY];
                nb[k][5] = Q[u + HX * RY];
            }
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            if ((uint32_t)(wave0 + NT * k) < n) {
                const uint32_t e = lc ? w[1][k] : w[0][k];
                float sum = 0.0f;
#pragma unroll
                for (int q = 0; q < 6; q++) sum = sum + nb[k][q];
                const float v = (sum - (lc ? bt[1][k] : bt[0][k])) / (float)((e >> 14) & 7u);
                if ((uint32_t)(tid + NT * k) < n) Pc[e & 0x1fffu] = v;
            }
        }
        __syncthreads();
    }

    // the core
    constexpr int CS = (CX * CY * CZ + NT - 1) / NT;
#pragma unroll
    for (int m = 0; m < CS; m++) {
        const int q = tid + NT * m;
        const int x = q % CX, y = (q / CX) % CY, z = q / (CX * CY);
        const int gx = x0 + x, gy = y0 + y, gz = z0 + z;
        if (q < CX * CY * CZ && gx < nx && gy < ny && gz < nz) {
            const int rx = x + H, ry = y + H, rz = z + H;
            pout[(uint32_t)gx + (uint32_t)nx * ((uint32_t)gy + (uint32_t)ny * (uint32_t)gz)] =
                P[((rx + ry + rz) & 1) * NC + (ry + RY * rz) * HX + (rx >> 1)];
        }
    }
}

inline float *r_buf(const DevSim &S, uint32_t k) { return (k & 1) ? S.cg_r1 : S.cg_r0; }
inline float *w_buf(const DevSim &S, uint32_t k) { return (k & 1) ? S.cg_w1 : S.cg_w0; }
inline float *s_buf(const DevSim &S, uint32_t k) { return (k & 1) ? S.cg_s1 : S.cg_s0; }

}  // namespace

uint32_t solver_blocks(const DevSim &S) { return (uint32_t)S.cg_blocks; }

void launch_cg_start(const DevSim &S, bool fp64, float *x, bool inner, hipStream_t s) {
    const uint32_t g = solver_blocks(S);
    PXL_LAUNCH(k_cg_residual, g, SBLOCK, 0, s, S, x, r_buf(S, 0), inner ? 1 : 0);
    // iteration 0 reads s_{-1} from s_buf(1)
    if (fp64) PXL_LAUNCH((k_cg_start<double>), g, SBLOCK, 0, s, S, r_buf(S, 0), w_buf(S, 0), s_buf(S, 1));
    else PXL_LAUNCH((k_cg_start<float>), g, SBLOCK, 0, s, S, r_buf(S, 0), w_buf(S, 0), s_buf(S, 1));
}

void launch_cg_iter(const DevSim &S, bool fp64, float *x, uint32_t it, hipStream_t s) {
    const uint32_t g = solver_blocks(S);
    float *rin = r_buf(S, it), *rout = r_buf(S, it + 1), *win = w_buf(S, it), *wout = w_buf(S, it + 1);
    float *sprev = s_buf(S, it + 1), *scur = s_buf(S, it);
    if (fp64) PXL_LAUNCH((k_cg_iter<double>), g, SBLOCK, 0, s, S, it, x, rin, rout, win, wout, sprev, scur);
    else PXL_LAUNCH((k_cg_iter<float>), g, SBLOCK, 0, s, S, it, x, rin, rout, win, wout, sprev, scur);
}

void launch_mgcg_start(const DevSim &S, float *x, hipStream_t s) {
    const uint32_t g = solver_blocks(S);
    PXL_LAUNCH(k_cg_residual, g, SBLOCK, 0, s, S, x, S.cg_r0, 1);
    launch_mg_vcycle(S, S.cg_r0, S.cg_r1, s);
    PXL_LAUNCH((k_mgcg_spmv<double>), g, SBLOCK, 0, s, S, 0, 1);
}

void launch_mgcg_iter(const DevSim &S, float *x, uint32_t it, bool last, hipStream_t s) {
    const uint32_t g = solver_blocks(S);
    PXL_LAUNCH((k_mgcg_update<double>), g, SBLOCK, 0, s, S, it, last ? 1 : 0, x);
    if (last) return;
    launch_mg_vcycle(S, S.cg_r0, S.cg_r1, s);
    PXL_LAUNCH((k_mgcg_spmv<double>), g, SBLOCK, 0, s, S, (it + 1) & 1, 0);
}

void launch_mixed_init(const DevSim &S, double *x64, hipStream_t s) {
    PXL_LAUNCH(k_mixed_init, solver_blocks(S), SBLOCK, 0, s, S, x64);
    PXL_LAUNCH(k_norms_finalize, 1, BLOCK, 0, s, S, solver_blocks(S), 2);
}

void launch_mixed_update(const DevSim &S, const double *xsrc, double *xdst, hipStream_t s) {
    PXL_LAUNCH(k_mixed_update, solver_blocks(S), SBLOCK, 0, s, S, xsrc, xdst, S.e);
    PXL_LAUNCH(k_norms_finalize, 1, BLOCK, 0, s, S, solver_blocks(S), 1);
}


uint32_t rbgs_cores(const DevSim &S) {
    return (((uint32_t)S.nx + RB_CX - 1) / RB_CX) * (((uint32_t)S.ny + RB_CY - 1) / RB_CY) *
           (((uint32_t)S.nz + RB_CZ - 1) / RB_CZ);
}
uint32_t rbgs_list_slots() { return 2u * RbCore::LMAX; }

void launch_rbgs_blocked(const DevSim &S, uint32_t iters, hipStream_t s) {
    const uint32_t sweeps = 2 * iters;
    if (sweeps == 0) return;
    const uint32_t g = rbgs_cores(S);
    PXL_LAUNCH((k_rbgs_lists<RB_CX, RB_CY, RB_CZ, RB_NT>), g, RB_NT, 0, s, S);
    // an even number of launches (the last then writes the pressure array), each of at most
    // RB_HALO half sweeps
    uint32_t launches = (sweeps + RB_HALO - 1) / RB_HALO;
    launches += launches & 1u;
    const uint32_t base = sweeps / launches, extra = sweeps % launches;
    float *a = S.pressure, *b = S.cg_r0;  // the CG vectors are free in reference mode
    uint32_t done = 0;
    for (uint32_t l = 0; l < launches; l++) {
        const uint32_t n = base + (l < extra ? 1u : 0u);
        if (g <= RB_WIDE_CORES)
            PXL_LAUNCH((k_rbgs_tile<RB_CX, RB_CY, RB_CZ, RB_NT>), g, RB_NT, 0, s, S, a, b, done & 1u, n, l == 0 ? 1 : 0);
        else
            PXL_LAUNCH((k_rbgs_tile<RB_CX, RB_CY, RB_CZ, 512>), g, 512, 0, s, S, a, b, done & 1u, n, l == 0 ? 1 : 0);
        done += n;
        std::swap(a, b);
    }
}

// ---- batched (the solving simulations: x = e, r = cg_r0, u = cg_r1)
void batch_mixed_init(const WorkLists &W, hipStream_t s) { PXL_LAUNCH(kb_mixed_init, batch_grid(W[WD_A_SOLVER], batch_cap(BF_SOLVER)), SBLOCK, 0, s, W[WD_A_SOLVER]); }
void batch_norms(const WorkLists &W, int nv, hipStream_t s) {
    PXL_LAUNCH(kb_norms_finalize, batch_grid(W[WD_A_SIM], batch_cap(BF_SOLVER)), BLOCK, 0, s, W[WD_A_SIM], nv);
}
void batch_cg_residual(const WorkLists &W, hipStream_t s) { PXL_LAUNCH(kb_cg_residual, batch_grid(W[WD_A_SOLVER], batch_cap(BF_SOLVER)), SBLOCK, 0, s, W[WD_A_SOLVER]); }
void batch_mgcg_spmv(const WorkLists &W, uint32_t slot, int first, hipStream_t s) {
    PXL_LAUNCH(kb_mgcg_spmv, batch_grid(W[WD_A_SOLVER], batch_cap(BF_SOLVER)), SBLOCK, 0, s, W[WD_A_SOLVER], slot, first);
}
void batch_mgcg_update(const WorkLists &W, uint32_t it, int last, hipStream_t s) {
    PXL_LAUNCH(kb_mgcg_update, batch_grid(W[WD_A_SOLVER], batch_cap(BF_SOLVER)), SBLOCK, 0, s, W[WD_A_SOLVER], it, last);
}
void batch_mixed_update(const WorkLists &W, int cur, hipStream_t s) {
    PXL_LAUNCH(kb_mixed_update, batch_grid(W[WD_A_SOLVER], batch_cap(BF_SOLVER)), SBLOCK, 0, s, W[WD_A_SOLVER], cur);
}

}  // namespace pxl
