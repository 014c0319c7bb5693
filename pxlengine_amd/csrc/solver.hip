// Pressure solvers on the matrix of the reference's Gauss-Seidel sweep (fluid.slang:168-182):
// for FLUID cell i, (A p)_i = cnt_i p_i - sum_{FLUID nbrs} p_j (cnt_i = non-SOLID neighbours,
// AIR neighbours are Dirichlet 0), b_i = -div_i h^2 / dt.
//
// Jacobi-preconditioned CG, two fused passes per iteration:
//   pass A: d = z + beta*d_prev (z = r / cnt) at the cell and its FLUID neighbours,
//           q = A d, and the workgroup partial of d.q; the last workgroup forms alpha.
//   pass B: x += alpha d, r -= alpha q, and the partial of r.z; the last workgroup forms beta.
// The direction vector ping-pongs between two buffers so pass A can rebuild neighbours' d
// from unmodified inputs.  Accumulation type Acc = float (plain FP32 CG) or double (FP64
// dot products of the mixed-precision solve); vector arithmetic is fp32 in both.
//
// Mixed precision: the FP64 outer loop keeps x64 (solution) and r64 = b - A x64 in FP64,
// solves A e = float(r64) with the FP32 iterations above from e = 0, then x64 += e and
// recomputes r64 in FP64.
#include "common.hpp"
#include "kernels.hpp"

namespace pxl {

namespace {

constexpr int TX = 64, TY = 4;

struct CellPos {
    int i, j, k;
    uint32_t id;
    bool valid;
};

__device__ __forceinline__ CellPos cell_pos(const DevSim &S) {
    CellPos c;
    c.i = blockIdx.x * TX + threadIdx.x;
    c.j = blockIdx.y * TY + threadIdx.y;
    c.k = blockIdx.z;
    c.valid = c.i < S.nx && c.j < S.ny;
    c.id = (uint32_t)c.i + (uint32_t)S.nx * ((uint32_t)c.j + (uint32_t)S.ny * (uint32_t)c.k);
    return c;
}

__device__ __forceinline__ float diag_of(uint32_t st) { return (float)((st >> 8) & 7u); }

// z = r / cnt (0 where cnt = 0).
__device__ __forceinline__ float precond(float r, uint32_t st) {
    const float dg = diag_of(st);
    return dg > 0.0f ? r / dg : 0.0f;
}

// Offsets of the six neighbours, order -x,+x,-y,+y,-z,+z.
__device__ __forceinline__ void nb_offsets(const DevSim &S, int32_t off[6]) {
    const int32_t sy = S.nx, sz = S.nx * S.ny;
    off[0] = -1; off[1] = 1; off[2] = -sy; off[3] = sy; off[4] = -sz; off[5] = sz;
}

template <typename T>
__device__ __forceinline__ T to_acc(float v) { return (T)v; }

// ---- init: r = b - A x for the pressure system (x = warm-started pressure), rz = r.z
template <typename Acc>
__global__ __launch_bounds__(TX *TY) void k_cg_init(DevSim S, float *x, float *dprev) {
    const CellPos c = cell_pos(S);
    Acc rz = 0;
    if (c.valid) {
        const uint32_t st = S.sten[c.id];
        if (st & STEN_FLUID) {
            int32_t off[6];
            nb_offsets(S, off);
            float sum = 0.0f;
#pragma unroll
            for (int n = 0; n < 6; n++)
                if (st & (1u << n)) sum = sum + x[c.id + off[n]];
            const float ax = diag_of(st) * x[c.id] - sum;
            const float b = -(S.div[c.id] * S.h * S.h / S.dt);
            const float r = b - ax;
            S.r[c.id] = r;
            dprev[c.id] = 0.0f;
            if (diag_of(st) > 0.0f) rz = to_acc<Acc>(r) * to_acc<Acc>(r / diag_of(st));
        }
    }
    const Acc v[1] = {rz};
    Acc tot[1];
    if (grid_sum<Acc, 1>(v, reinterpret_cast<Acc *>(sizeof(Acc) == 8 ? (void *)S.part64 : (void *)S.part32), S.ctr, tot) &&
        flat_tid() == 0) {
        S.scal[SC_RZ] = (double)tot[0];
        S.scal[SC_BETA] = 0.0;
    }
}

// ---- mixed inner init: r = float(r64), e = 0, rz = r.z (FP64)
__global__ __launch_bounds__(TX *TY) void k_inner_init(DevSim S, float *e, float *dprev) {
    const CellPos c = cell_pos(S);
    double rz = 0;
    if (c.valid) {
        const uint32_t st = S.sten[c.id];
        if (st & STEN_FLUID) {
            const float r = (float)S.r64[c.id];
            S.r[c.id] = r;
            e[c.id] = 0.0f;
            dprev[c.id] = 0.0f;
            if (diag_of(st) > 0.0f) rz = (double)r * (double)(r / diag_of(st));
        }
    }
    const double v[1] = {rz};
    double tot[1];
    if (grid_sum<double, 1>(v, S.part64, S.ctr, tot) && flat_tid() == 0) {
        S.scal[SC_RZ] = tot[0];
        S.scal[SC_BETA] = 0.0;
    }
}

// ---- pass A: d = z + beta d_prev; q = A d; alpha = rz / (d.q)
template <typename Acc>
__global__ __launch_bounds__(TX *TY) void k_cg_dir_spmv(DevSim S, const float *__restrict__ dprev,
                                                       float *__restrict__ dcur) {
    const CellPos c = cell_pos(S);
    const float beta = (float)S.scal[SC_BETA];
    Acc dq = 0;
    if (c.valid) {
        const uint32_t st = S.sten[c.id];
        if (st & STEN_FLUID) {
            int32_t off[6];
            nb_offsets(S, off);
            const float dc = precond(S.r[c.id], st) + beta * dprev[c.id];
            float sum = 0.0f;
#pragma unroll
            for (int n = 0; n < 6; n++) {
                if (st & (1u << n)) {
                    const uint32_t nid = c.id + off[n];
                    const float dn = precond(S.r[nid], S.sten[nid]) + beta * dprev[nid];
                    sum = sum + dn;
                }
            }
            const float qc = diag_of(st) * dc - sum;
            dcur[c.id] = dc;
            S.q[c.id] = qc;
            dq = to_acc<Acc>(dc) * to_acc<Acc>(qc);
        }
    }
    const Acc v[1] = {dq};
    Acc tot[1];
    if (grid_sum<Acc, 1>(v, reinterpret_cast<Acc *>(sizeof(Acc) == 8 ? (void *)S.part64 : (void *)S.part32), S.ctr, tot) &&
        flat_tid() == 0) {
        const Acc rz = (Acc)S.scal[SC_RZ];
        S.scal[SC_ALPHA] = (tot[0] != (Acc)0 && rz != (Acc)0) ? (double)(rz / tot[0]) : 0.0;
    }
}

// ---- pass B: x += alpha d; r -= alpha q; beta = rz_new / rz
template <typename Acc>
__global__ __launch_bounds__(TX *TY) void k_cg_update(DevSim S, float *__restrict__ x, const float *__restrict__ d) {
    const CellPos c = cell_pos(S);
    const float alpha = (float)S.scal[SC_ALPHA];
    Acc rz = 0;
    if (c.valid) {
        const uint32_t st = S.sten[c.id];
        if (st & STEN_FLUID) {
            x[c.id] = x[c.id] + alpha * d[c.id];
            const float r = S.r[c.id] - alpha * S.q[c.id];
            S.r[c.id] = r;
            if (diag_of(st) > 0.0f) rz = to_acc<Acc>(r) * to_acc<Acc>(r / diag_of(st));
        }
    }
    const Acc v[1] = {rz};
    Acc tot[1];
    if (grid_sum<Acc, 1>(v, reinterpret_cast<Acc *>(sizeof(Acc) == 8 ? (void *)S.part64 : (void *)S.part32), S.ctr, tot) &&
        flat_tid() == 0) {
        const Acc rz_old = (Acc)S.scal[SC_RZ];
        S.scal[SC_BETA] = rz_old != (Acc)0 ? (double)(tot[0] / rz_old) : 0.0;
        S.scal[SC_RZ] = (double)tot[0];
    }
}

// ---- FP64 outer loop.  The solution is double-buffered (x_src -> x_dst) so the update pass
// can rebuild neighbours' x_src + e from values no lane has overwritten.
__device__ __forceinline__ double b64_of(const DevSim &S, uint32_t id) {
    return -((double)S.div[id] * (double)S.h * (double)S.h / (double)S.dt);
}

// init: x64 = pressure, r64 = b - A x64; totals rr and bb.
__global__ __launch_bounds__(TX *TY) void k_mixed_init(DevSim S, double *__restrict__ x64) {
    const CellPos c = cell_pos(S);
    double rr = 0, bb = 0;
    if (c.valid) {
        const uint32_t st = S.sten[c.id];
        if (st & STEN_FLUID) {
            int32_t off[6];
            nb_offsets(S, off);
            double sum = 0.0;
#pragma unroll
            for (int n = 0; n < 6; n++)
                if (st & (1u << n)) sum += (double)S.pressure[c.id + off[n]];
            const double xc = (double)S.pressure[c.id];
            const double b = b64_of(S, c.id);
            const double r = b - ((double)diag_of(st) * xc - sum);
            x64[c.id] = xc;
            S.r64[c.id] = r;
            rr = r * r;
            bb = b * b;
        }
    }
    const double v[2] = {rr, bb};
    double tot[2];
    if (grid_sum<double, 2>(v, S.part64, S.ctr, tot) && flat_tid() == 0) {
        S.scal[SC_RR] = tot[0];
        S.scal[SC_BB] = tot[1];
    }
}

// update: x_dst = x_src + e (FP64) at the cell and its neighbours, r64 = b - A x_dst; rr.
__global__ __launch_bounds__(TX *TY) void k_mixed_update(DevSim S, const double *__restrict__ xsrc,
                                                         double *__restrict__ xdst, const float *__restrict__ e) {
    const CellPos c = cell_pos(S);
    double rr = 0;
    if (c.valid) {
        const uint32_t st = S.sten[c.id];
        if (st & STEN_FLUID) {
            int32_t off[6];
            nb_offsets(S, off);
            double sum = 0.0;
#pragma unroll
            for (int n = 0; n < 6; n++) {
                if (st & (1u << n)) {
                    const uint32_t nid = c.id + off[n];
                    sum += xsrc[nid] + (double)e[nid];
                }
            }
            const double xc = xsrc[c.id] + (double)e[c.id];
            const double r = b64_of(S, c.id) - ((double)diag_of(st) * xc - sum);
            xdst[c.id] = xc;
            S.r64[c.id] = r;
            rr = r * r;
        }
    }
    const double v[1] = {rr};
    double tot[1];
    if (grid_sum<double, 1>(v, S.part64, S.ctr, tot) && flat_tid() == 0) S.scal[SC_RR] = tot[0];
}

__global__ __launch_bounds__(TX *TY) void k_mixed_finish(DevSim S, const double *x64) {
    const CellPos c = cell_pos(S);
    if (c.valid && (S.sten[c.id] & STEN_FLUID)) S.pressure[c.id] = (float)x64[c.id];
}

dim3 cgrid(const DevSim &S) { return dim3((S.nx + TX - 1) / TX, (S.ny + TY - 1) / TY, S.nz); }

}  // namespace

void launch_cg_init32(const DevSim &S, hipStream_t s) {
    k_cg_init<float><<<cgrid(S), dim3(TX, TY), 0, s>>>(S, S.pressure, S.d1);
}

void launch_inner_init(const DevSim &S, hipStream_t s) {
    k_inner_init<<<cgrid(S), dim3(TX, TY), 0, s>>>(S, S.e, S.d1);
}

// Iteration `it` reads the direction of iteration it-1 from d1 (even it) or d0 (odd it).
void launch_cg_pass_a(const DevSim &S, bool fp64, uint32_t it, hipStream_t s) {
    const float *dprev = (it & 1) ? S.d0 : S.d1;
    float *dcur = (it & 1) ? S.d1 : S.d0;
    if (fp64) k_cg_dir_spmv<double><<<cgrid(S), dim3(TX, TY), 0, s>>>(S, dprev, dcur);
    else k_cg_dir_spmv<float><<<cgrid(S), dim3(TX, TY), 0, s>>>(S, dprev, dcur);
}

void launch_cg_pass_b(const DevSim &S, bool fp64, float *x, uint32_t it, hipStream_t s) {
    const float *dcur = (it & 1) ? S.d1 : S.d0;
    if (fp64) k_cg_update<double><<<cgrid(S), dim3(TX, TY), 0, s>>>(S, x, dcur);
    else k_cg_update<float><<<cgrid(S), dim3(TX, TY), 0, s>>>(S, x, dcur);
}

void launch_mixed_init(const DevSim &S, double *x64, hipStream_t s) {
    k_mixed_init<<<cgrid(S), dim3(TX, TY), 0, s>>>(S, x64);
}

void launch_mixed_update(const DevSim &S, const double *xsrc, double *xdst, hipStream_t s) {
    k_mixed_update<<<cgrid(S), dim3(TX, TY), 0, s>>>(S, xsrc, xdst, S.e);
}

void launch_mixed_finish(const DevSim &S, const double *x64, hipStream_t s) {
    k_mixed_finish<<<cgrid(S), dim3(TX, TY), 0, s>>>(S, x64);
}

}  // namespace pxl
