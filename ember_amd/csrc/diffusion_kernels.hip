// Batched diffusion stage: all variables' tridiagonal systems of one grid.
//
// Replaces FlameSolver::integrateDiffusionTerms (src/flameSolver.cpp:838-854):
// per variable, TridiagonalIntegrator::integrateToTime (src/integrator.cpp:37-47)
// takes nSteps steps; the first is 8 BDF1 substeps of h/8, later steps are BDF2
// (src/integrator.cpp:133-204).  The matrix is DiffusionSystem::get_A / get_k
// (src/diffusionSystem.cpp:10-76) on OneDimGrid metrics (src/grid.cpp:43-69),
// evaluated on the fly from B, D and x instead of being stored.
//
// Layout: Ember's point-major state (nVars values of one point contiguous).  A
// wavefront owns a chunk of consecutive points, lane v = variable v, so every row
// load is one contiguous nVars-double segment.  Each linear solve is a partition
// method: the last point of every chunk is a separator; a chunk eliminates its
// interior towards both separators (downward and upward sweeps), the separators
// form a reduced tridiagonal system, which is solved by the same scheme one level
// up, and finally every chunk solves its interior against known separators.  The fused
// stage path (below) takes the second level inside the chunk kernel's workgroups and
// solves the remaining system of one row per workgroup with one wave per variable.
#include <hip/hip_runtime.h>

#include "emx_common.h"
#include "kernels.h"
#include "wave.h"

#include <algorithm>

namespace emx {

namespace {

// ---------------------------------------------------------------------------
// level-1 rows: shifted diffusion matrix and the BDF right-hand side
// ---------------------------------------------------------------------------
struct DiffRows {
    const double* B;
    const int32_t* B_row;
    const double* D;
    const double* sc;
    const double* x;
    const double* y;
    const double* yprev;
    const double* h;      // per variable step
    int N, nV, alpha, left, right;
    double yInf, wallConst;
    int wallVar;          // variable carrying yInf / wallConst (others: 0)
    int q;                // solve index: 0..7 BDF1 substeps, >= 8 BDF2 steps
    int bdf2;             // fused stage: it has BDF2 solves (solve 0 keeps y_0 for the first)

    __device__ __forceinline__ double Bv(int v, int j) const {
        return B[(size_t)(B_row ? B_row[v] : v) * N + j];
    }
    __device__ __forceinline__ double Dv(int v, int j) const { return D[(size_t)j * nV + v]; }
    __device__ __forceinline__ double c1(int v, int j) const {
        if (j < 1 || j > N - 2) return 0.0;
        double dlj = 0.5 * (x[j + 1] - x[j - 1]);
        double r = alpha ? x[j] : 1.0;
        return 0.5 * Bv(v, j) / (dlj * r);
    }
    __device__ __forceinline__ double c2(int v, int j) const {
        if (j < 1 || j > N - 2) return 0.0;
        double hh = x[j + 1] - x[j];
        double rph = alpha ? 0.5 * (x[j] + x[j + 1]) : 1.0;
        return rph * (Dv(v, j) + Dv(v, j + 1)) / hh;
    }
    // DiffusionSystem::get_A row j and get_k
    __device__ void matrix(int v, int j, double& a, double& b, double& c, double& k) const {
        a = 0.0;
        b = 0.0;
        c = 0.0;
        const int jStart = (left == 1) ? 2 : 1;
        const int jStop = (right == 1) ? N - 2 : N - 1;
        if (j == 0) {
            if (left == 2) {  // ControlVolume
                double hh0 = x[1] - x[0];
                double c0 = Bv(v, 0) * (alpha + 1) * (Dv(v, 0) + Dv(v, 1)) / (2 * hh0 * hh0);
                b = -c0;
                c = c0;
            } else if (left == 3) {  // WallFlux
                double hh0 = x[1] - x[0];
                double c0 = Bv(v, 0) * (alpha + 1) / hh0;
                double d = 0.5 * (Dv(v, 0) + Dv(v, 1));
                b = -c0 * (d / hh0 + (v == wallVar ? wallConst : 0.0));
                c = d * c0 / hh0;
            }
        } else if (left == 1 && j == 1) {  // ZeroGradient (left)
            double cc = c1(v, 1) * c2(v, 1);
            b = -cc;
            c = cc;
        } else if (right == 1 && j == N - 2) {  // ZeroGradient (right)
            double cc = c1(v, N - 2) * c2(v, N - 3);
            a = cc;
            b = -cc;
        } else if (j >= jStart && j < jStop) {
            double c1j = c1(v, j), c2m = c2(v, j - 1), c2j = c2(v, j);
            a = c1j * c2m;
            b = -c1j * (c2m + c2j);
            c = c1j * c2j;
        }
        k = sc ? sc[(size_t)j * nV + v] : 0.0;
        if (j == 0 && left == 3) {
            double hh0 = x[1] - x[0];
            k += (v == wallVar) ? Bv(v, 0) * (alpha + 1) / hh0 * wallConst * yInf : 0.0;
        }
    }
    __device__ __forceinline__ void row(int v, int j, double& a, double& b, double& c,
                                        double& r) const {
        double k;
        matrix(v, j, a, b, c, k);
        const double hv = h[v];
        const double yj = y[(size_t)j * nV + v];
        if (q < 8) {
            b = b - 1.0 / (hv / 8.0);
            r = -yj / (hv / 8.0) - k;
        } else {
            b = b - 3.0 / (2.0 * hv);
            r = -2 * yj / hv + yprev[(size_t)j * nV + v] / (2 * hv) - k;
        }
    }
};

// explicit (a, b, c, k) systems, system-major (test/tridiagonalIntegrator.cpp)
struct ExplicitRows {
    const double *a, *b, *c, *k, *y, *yprev, *h;
    int N, nV, q;
    __device__ __forceinline__ void row(int v, int j, double& aa, double& bb, double& cc,
                                        double& r) const {
        size_t o = (size_t)v * N + j;
        aa = a[o];
        bb = b[o];
        cc = c[o];
        const double hv = h[v];
        const double yj = y[(size_t)j * nV + v];
        if (q < 8) {
            bb = bb - 1.0 / (hv / 8.0);
            r = -yj / (hv / 8.0) - k[o];
        } else {
            bb = bb - 3.0 / (2.0 * hv);
            r = -2 * yj / hv + yprev[(size_t)j * nV + v] / (2 * hv) - k[o];
        }
    }
};

// reduced rows assembled from the chunk summaries of the level below:
// own[c] = (alpha_c, beta_c, rho_c), nxt[c] = (dbeta, gamma, drho) for row c-1
struct ReducedRows {
    const double* own;
    const double* nxt;
    int C, nV;
    __device__ __forceinline__ void row(int v, int c, double& a, double& b, double& cc,
                                        double& r) const {
        const double* o = own + ((size_t)c * nV + v) * 3;
        a = o[0];
        b = o[1];
        r = o[2];
        cc = 0.0;
        if (c + 1 < C) {
            const double* n = nxt + ((size_t)(c + 1) * nV + v) * 3;
            b += n[0];
            cc = n[1];
            r += n[2];
        }
    }
};

// ---------------------------------------------------------------------------
// chunk elimination: summary of chunk [s, e] (e = separator) for variable v
// ---------------------------------------------------------------------------
template <class Rows>
__device__ void chunk_summary(const Rows& R, int v, int s, int e, double* own, double* nxt) {
    double bt = 1.0, rt = 0.0, at = 0.0, cL = 0.0;
    // downward sweep over the interior s..e-1
    for (int i = s; i < e; i++) {
        double a, b, c, r;
        R.row(v, i, a, b, c, r);
        if (i == s) {
            bt = b;
            rt = r;
            at = a;
        } else {
            double f = a / bt;
            bt = b - f * cL;
            rt = r - f * rt;
            at = -f * at;
        }
        cL = c;
    }
    // upward sweep over the interior e-1..s
    double bh = 1.0, rh = 0.0, ch = 0.0, aN = 0.0;
    for (int i = e - 1; i >= s; i--) {
        double a, b, c, r;
        R.row(v, i, a, b, c, r);
        if (i == e - 1) {
            bh = b;
            rh = r;
            ch = c;
        } else {
            double g = c / bh;
            bh = b - g * aN;
            rh = r - g * rh;
            ch = -g * ch;
        }
        aN = a;
    }
    double ae, be, ce, re;
    R.row(v, e, ae, be, ce, re);
    (void)ce;
    const int m = e - s;
    if (m > 0) {
        own[0] = -ae * at / bt;
        own[1] = be - ae * cL / bt;
        own[2] = re - ae * rt / bt;
    } else {
        own[0] = ae;
        own[1] = be;
        own[2] = re;
    }
    if (s > 0) {
        double ap, bp, cp, rp;
        R.row(v, s - 1, ap, bp, cp, rp);
        if (m > 0) {
            nxt[0] = -cp * aN / bh;
            nxt[1] = -cp * ch / bh;
            nxt[2] = -cp * rh / bh;
        } else {
            nxt[0] = 0.0;
            nxt[1] = cp;
            nxt[2] = 0.0;
        }
    } else {
        nxt[0] = nxt[1] = nxt[2] = 0.0;
    }
}

// Thomas over the interior s..e-1 with x_{s-1} = xL, x_e = xR known.
template <int P, class Rows, class Out>
__device__ void chunk_solve(const Rows& R, int v, int s, int e, double xL, double xR, Out& out) {
    double bt[P], rt[P], cc[P];
    const int m = e - s;
#pragma unroll
    for (int t = 0; t < P; t++) {
        if (t < m) {
            int i = s + t;
            double a, b, c, r;
            R.row(v, i, a, b, c, r);
            if (t == 0) r -= a * xL;
            if (t == m - 1) r -= c * xR;
            if (t == 0) {
                bt[t] = b;
                rt[t] = r;
            } else {
                double f = a / bt[t - 1];
                bt[t] = b - f * cc[t - 1];
                rt[t] = r - f * rt[t - 1];
            }
            cc[t] = c;
        }
    }
    double xn = xR;
#pragma unroll
    for (int t = P - 1; t >= 0; t--) {
        if (t < m) {
            double xi = (t == m - 1) ? rt[t] / bt[t] : (rt[t] - cc[t] * xn) / bt[t];
            out.put(v, s + t, xi);
            xn = xi;
        }
    }
    out.put(v, e, xR);
}

struct StateOut {
    double* y;
    double* yprev;
    int nV;
    bool save_prev;
    __device__ __forceinline__ void put(int v, int j, double val) const {
        size_t o = (size_t)j * nV + v;
        if (save_prev) yprev[o] = y[o];
        y[o] = val;
    }
};

struct SepOut {
    double* X;
    int nV;
    __device__ __forceinline__ void put(int v, int c, double val) const { X[(size_t)c * nV + v] = val; }
};

__device__ __forceinline__ bool var_active(const int* nsteps, int v, int q) {
    return !nsteps || q < 8 || (q - 7) < nsteps[v];
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <int P, class Rows>
__global__ void __launch_bounds__(256) summary_kernel(Rows R, int n, int nV, const int* nsteps,
                                                      int q, double* own, double* nxt) {
    const int lane = lane_id();
    const int chunk = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int C = (n + P - 1) / P;
    if (chunk >= C || lane >= nV) return;
    if (nsteps && !var_active(nsteps, lane, q)) return;
    const int s = chunk * P;
    const int e = min(s + P, n) - 1;
    chunk_summary(R, lane, s, e, own + ((size_t)chunk * nV + lane) * 3,
                  nxt + ((size_t)chunk * nV + lane) * 3);
}

template <int P, class Rows, class Out>
__global__ void __launch_bounds__(256) solve_kernel(Rows R, Out out, int n, int nV,
                                                    const int* nsteps, int q, const double* X) {
    const int lane = lane_id();
    const int chunk = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int C = (n + P - 1) / P;
    if (chunk >= C || lane >= nV) return;
    if (nsteps && !var_active(nsteps, lane, q)) return;
    const int s = chunk * P;
    const int e = min(s + P, n) - 1;
    const double xL = chunk > 0 ? X[(size_t)(chunk - 1) * nV + lane] : 0.0;
    const double xR = X[(size_t)chunk * nV + lane];
    chunk_solve<P>(R, lane, s, e, xL, xR, out);
}

// Sequential Thomas over all rows of the top-level system, one lane per variable.
template <class Rows, class Out>
__global__ void thomas_kernel(Rows R, Out out, int n, int nV, const int* nsteps, int q,
                              double* cp, double* dp) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    if (nsteps && !var_active(nsteps, v, q)) return;
    double bprev = 1.0, dprev = 0.0, cprev = 0.0;
    for (int i = 0; i < n; i++) {
        double a, b, c, r;
        R.row(v, i, a, b, c, r);
        double den = (i == 0) ? b : b - a * cprev;
        double inv = 1.0 / den;
        double ci = c * inv;
        double di = (i == 0) ? r * inv : (r - a * dprev) * inv;
        cp[(size_t)i * nV + v] = ci;
        dp[(size_t)i * nV + v] = di;
        cprev = ci;
        dprev = di;
        bprev = den;
    }
    (void)bprev;
    double xn = 0.0;
    for (int i = n - 1; i >= 0; i--) {
        double xi = (i == n - 1) ? dp[(size_t)i * nV + v] : dp[(size_t)i * nV + v] - cp[(size_t)i * nV + v] * xn;
        out.put(v, i, xi);
        xn = xi;
    }
}

// Ember's diffusion step rule: per chunk min over j in [1, N-2] of dlj^2 / (B D)
__global__ void __launch_bounds__(256) hmin_chunk_kernel(DiffRows R, int P, double* cmin) {
    const int lane = lane_id();
    const int chunk = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int C = (R.N + P - 1) / P;
    if (chunk >= C || lane >= R.nV) return;
    double mn = __builtin_huge_val();
    const int s = chunk * P, e = min(s + P, R.N);
    for (int j = max(s, 1); j < min(e, R.N - 1); j++) {
        double dl = 0.5 * (R.x[j + 1] - R.x[j - 1]);
        mn = fmin(mn, dl * dl / (R.Bv(lane, j) * R.Dv(lane, j)));
    }
    cmin[(size_t)chunk * R.nV + lane] = mn;
}

__global__ void hmin_reduce_kernel(const double* cmin, int C, int nV, double* out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    double mn = __builtin_huge_val();
    for (int c = 0; c < C; c++) mn = fmin(mn, cmin[(size_t)c * nV + v]);
    out[v] = mn;
}

// ---------------------------------------------------------------------------
// Fused stage path (C1 = ceil(N / P1) <= kPcrMax): one setup launch per stage (grid
// metrics, edge-chunk rows, step sizes), then per solve one separator launch (pcr_kernel)
// and one chunk launch that back-substitutes this solve and eliminates the chunks for the
// next solve with the matrix rows still in registers; a workgroup's chunks form one
// workgroup chunk (two-level partition), so the reduced system has G = ceil(C1 / 4) rows.
//
// The reduced system of this path is variable-major: plane k of the workgroup summaries
// (own: alpha, beta, rho; nxt: dbeta, gamma, drho) holds element (v, g) at
// (k * nV + v) * G + g, and the separators X at v * G + g, so the separator workgroup of
// one variable reads and writes contiguous rows.
// ---------------------------------------------------------------------------
constexpr int kPcrMax = 4096;
// buffer resource word 3 of gfx9 (raw 32-bit records, no swizzle)
constexpr int kBufferFlags = 0x00020000;
// waves per workgroup of diff_step_kernel
constexpr int kDiffWaves = 4;

// 1 / b by Newton-Raphson from the hardware estimate (v_rcp_f64, relative error below
// 2^-22): each step r <- r + r (1 - b r) squares the relative error, so two steps reach
// the rounding level.  For the pivots of diagonally dominant rows (no zero, overflow or
// denormal range).
__device__ __forceinline__ double rcp_nr(double b) {
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    e = fma(-b, r, 1.0);
    return fma(r, e, r);
}

// TridiagonalIntegrator::step rows of solve q (src/integrator.cpp:133-204): diagonal shifted
// by sh and right-hand side cy y + cp yp - k, with BDF1 substeps of h/8 (q < 8:
// sh = 8/h, cy = -8/h, cp = 0) and BDF2 steps of h (sh = 3/(2h), cy = -2/h, cp = 1/(2h)).
// The reference divides by h/8 and h; the products with the reciprocals agree to the
// last bit or one unit of it.
struct BdfCoef {
    double sh, cy, cp;
    __device__ __forceinline__ BdfCoef(int q, double hv) {
        const double ih = 1.0 / hv;
        if (q < 8) {
            sh = 8.0 * ih;
            cy = -sh;
            cp = 0.0;
        } else {
            sh = 1.5 * ih;
            cy = -2.0 * ih;
            cp = 0.5 * ih;
        }
    }
    __device__ __forceinline__ double rhs(double y, double yp, double k) const {
        return cy * y + cp * yp - k;
    }
};

// Point-major array (nV values of one point contiguous) read and written through a buffer
// resource: lane v supplies the byte offset 8 v, the row offset j * nV * 8 is a scalar, so
// all rows of a chunk share one address register.  Out-of-range accesses read 0 and
// write nothing; a null array has no records and reads as 0.
struct PtBuf {
    __amdgpu_buffer_rsrc_t r;
    unsigned rowb;  // bytes per point row
    __device__ __forceinline__ PtBuf(const double* p, int N, int nV)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(p), 0,
                                              p ? (int)((size_t)N * nV * sizeof(double)) : 0,
                                              kBufferFlags)),
          rowb((unsigned)nV * sizeof(double)) {}
    using u32x2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0u, 0u, 0));
    __device__ __forceinline__ double ld(unsigned v8, int j) const {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, v8, j * rowb, 0));
    }
    __device__ __forceinline__ void st(double x, unsigned v8, int j) const {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, v8, j * rowb, 0);
    }
};

// Edge chunks: the chunks that are not all generic interior rows (chunk 0, and the one or
// two chunks holding rows N-2 and N-1: chunk c is interior when s = cP >= 2 and
// s + P - 1 <= N - 3, so only chunks 0, C-2 and C-1 can fail it).  Their matrix rows
// (DiffusionSystem::get_A with the boundary conditions), right-hand-side constants (splitConst
// plus the WallFlux term of get_k) and the c coefficient of the row before the chunk depend
// on the grid, B, D and splitConst only, which a stage holds fixed, so they are formed once
// per stage (stage_setup_kernel) into a table that the edge chunks of every solve read with
// one batch of loads.  Table: slot e = 0, 1, 2 for chunks 0, C-2, C-1; part k = a, b, c, K
// of row t at ((e * 4 + k) * P + t) * 64 + v; cp of slot e at 3 * 4 * P * 64 + e * 64 + v.
__host__ __device__ __forceinline__ int edge_slot_chunk(int e, int C) {
    return e == 0 ? 0 : (e == 1 ? C - 2 : C - 1);
}
__device__ __forceinline__ int edge_slot(int chunk, int C) {
    return chunk == 0 ? 0 : (chunk == C - 1 ? 2 : 1);
}
template <int P>
constexpr int edge_doubles() { return 3 * 4 * P * 64 + 3 * 64; }

// One thread per (slot, row t in 0..P, variable): rows t < P of the chunk, and t = P for cp.
template <int P>
__device__ void edge_rows(const DiffRows& R, int idx, double* edge) {
    const int e = idx / ((P + 1) * 64), t = (idx / 64) % (P + 1), v = idx % 64;
    const int N = R.N, C = (N + P - 1) / P;
    if (e >= 3 || v >= R.nV) return;
    const int chunk = edge_slot_chunk(e, C);
    if (chunk < 0 || (e == 1 && chunk == 0)) return;  // C = 2: slot 1 would repeat chunk 0
    const int s = chunk * P, m = min(s + P, N) - s;
    if (t == P) {
        double a, b, c = 0.0, k;
        if (s > 0) R.matrix(v, s - 1, a, b, c, k);
        edge[3 * 4 * P * 64 + e * 64 + v] = c;
        return;
    }
    double a = 0.0, b = 0.0, c = 0.0, kk = 0.0;
    if (t < m) {
        const int j = s + t;
        double k;
        R.matrix(v, j, a, b, c, k);
        kk = R.sc ? R.sc[(size_t)j * R.nV + v] : 0.0;
        if (j == 0 && R.left == 3 && v == R.wallVar)  // WallFlux constant of get_k
            kk += R.Bv(v, 0) * (R.alpha + 1) / (R.x[1] - R.x[0]) * R.wallConst * R.yInf;
    }
    double* E = edge + (size_t)e * 4 * P * 64 + t * 64 + v;
    E[0] = a;
    E[P * 64] = b;
    E[2 * P * 64] = c;
    E[3 * P * 64] = kk;
}

// Once per stage: the grid metrics of OneDimGrid (src/grid.cpp:43-69) as they enter
// DiffusionSystem::get_A (src/diffusionSystem.cpp:10-50), c1_j = B_j g1_j with
// g1_j = 0.5 / (dlj r_j) and c2_j = (D_j + D_{j+1}) g2_j with g2_j = rphalf_j / hh_j (they
// depend on the grid only, so they are formed once per stage instead of once per variable
// and solve), in blocks [0, nbm); the edge table in the blocks after.
template <int P>
__global__ void __launch_bounds__(256) stage_setup_kernel(DiffRows R, int nbm, double* g1, double* g2,
                                                          double* edge, DiffSteps steps, double* h,
                                                          int* nsteps) {
    // the per-variable step sizes and step counts, delivered as a kernel argument so that the
    // values are fixed at launch (no host buffer has to outlive the call)
    if (blockIdx.x == 0 && (int)threadIdx.x < R.nV) {
        h[threadIdx.x] = steps.h[threadIdx.x];
        nsteps[threadIdx.x] = steps.ns[threadIdx.x];
    }
    if ((int)blockIdx.x >= nbm) {
        edge_rows<P>(R, ((int)blockIdx.x - nbm) * blockDim.x + threadIdx.x, edge);
        return;
    }
    const double* x = R.x;
    const int N = R.N, alpha = R.alpha;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    double a = 0.0, b = 0.0;
    if (j >= 1 && j <= N - 2) {
        const double dlj = 0.5 * (x[j + 1] - x[j - 1]);
        a = 0.5 / (dlj * (alpha ? x[j] : 1.0));
    }
    if (j <= N - 2) b = (alpha ? 0.5 * (x[j] + x[j + 1]) : 1.0) / (x[j + 1] - x[j]);
    g1[j] = a;
    g2[j] = b;
}

// Downward pivots of the chunk interior rows 0 .. m-2 for the shift sh: T1[t] = 1 / bt_t
// with bt_0 = b_0 - sh, bt_t = b_t - sh - a_t c_(t-1) / bt_(t-1).  The Thomas forward
// sweep of a solve and the downward sweep of a chunk summary share them.
template <int P, bool FULL>
__device__ __forceinline__ void down_pivots(int m, double sh, const double (&A)[P],
                                            const double (&Bm)[P], const double (&Cm)[P],
                                            double (&T1)[P]) {
#pragma unroll
    for (int t = 0; t < P - 1; t++) {
        if (FULL || t < m - 1) {
            const double bs = Bm[t] - sh;
            if (t == 0)
                T1[t] = rcp_nr(bs);
            else
                T1[t] = rcp_nr(bs - A[t] * T1[t - 1] * Cm[t - 1]);
        }
    }
}

// Two-level partition of the fused stage path.  The kDiffWaves chunks of a workgroup
// (chunks 4g .. 4g + kc - 1, kc <= 4) form one workgroup chunk whose separator is the last
// chunk's; the separators u_0 .. u_(kc-2) of its other chunks are eliminated inside the
// workgroup (the chunk summaries meet in LDS), so the reduced system that the separator
// kernel solves has one row per workgroup (G = ceil(C / 4) rows).  For the back-substitution
// the workgroup keeps, per interior separator j, the downward-sweep row
// bt_j u_j + at_j u_(-1) + c_j u_(j+1) = rt_j; u_(-1) and u_(kc-1) are the workgroup
// separators X2[g-1] (0 for g = 0) and X2[g].  Only rt_j depends on the solve's right-hand
// side: at_j, c_j and 1 / bt_j depend on the matrix and the BDF shift, which a stage holds
// fixed within each shift class (BDF1 substeps, BDF2 steps), so they are written by the
// first solve of each class only.  coef layout: rt at (g * 3 + j) * 64 + v; then, per class
// cl, (at, c, 1 / bt) of row j at 3 * 64 * G + ((cl * G + g) * 3 + j) * 3 * 64 + i * 64 + v.
constexpr int kWgRows = kDiffWaves - 1;  // interior separators per workgroup
__host__ __device__ __forceinline__ size_t wg_coef_doubles(int G) { return (size_t)(kWgRows + 2 * kWgRows * 3) * 64 * G; }
__device__ __forceinline__ double* wg_rhs(double* coef, int g, int v) { return coef + (size_t)g * kWgRows * 64 + v; }
__device__ __forceinline__ double* wg_mat(double* coef, int G, int cl, int g, int v) {
    return coef + (size_t)kWgRows * 64 * G + ((size_t)cl * G + g) * kWgRows * 3 * 64 + v;
}
struct WgSep {
    double UL, UR, cf[kWgRows][4];
    __device__ __forceinline__ void load(const double* X2, const double* coef, int v, int G, int g,
                                         int kc, int cl) {
        const double* Xv = X2 + (size_t)v * G;
        UL = g > 0 ? Xv[g - 1] : 0.0;
        UR = Xv[g];
        const double* cr = wg_rhs(const_cast<double*>(coef), g, v);
        const double* cm = wg_mat(const_cast<double*>(coef), G, cl, g, v);
#pragma unroll
        for (int j = 0; j < kWgRows; j++) {
            cf[j][0] = j < kc - 1 ? cr[j * 64] : 0.0;
#pragma unroll
            for (int i = 0; i < 3; i++) cf[j][i + 1] = j < kc - 1 ? cm[(j * 3 + i) * 64] : 0.0;
        }
    }
    // the separators left and right of chunk 4g + w
    __device__ __forceinline__ void get(int w, int kc, double& xL, double& xR) const {
        double u[kDiffWaves];
        u[kDiffWaves - 1] = UR;
#pragma unroll
        for (int j = kWgRows - 1; j >= 0; j--)
            u[j] = j < kc - 1 ? (cf[j][0] - cf[j][1] * UL - cf[j][2] * u[j + 1]) * cf[j][3] : UR;
        xR = w == 0 ? u[0] : (w == 1 ? u[1] : (w == 2 ? u[2] : u[3]));
        xL = w == 0 ? UL : (w == 1 ? u[0] : (w == 2 ? u[1] : u[2]));
    }
};

// Edge chunk (rows from the edge table; FULL: m == P, so no row needs a guard).
template <int P, bool FULL>
__device__ __forceinline__ bool chunk_step(const DiffRows& R, const double* __restrict__ edge,
                                           const int* nsteps, int q_solve, int q_next,
                                           const double* X2, const double* coef, int G, int kc,
                                           double* K, int v, int chunk, int C, double (&sm)[6]) {
    const int N = R.N, nV = R.nV;
    const int s = chunk * P;
    const int m = FULL ? P : min(s + P, N) - s;  // rows in the chunk; the last is the separator
    const unsigned v8 = (unsigned)v * sizeof(double);
    const PtBuf Yb(R.y, N, nV), Pb(R.yprev, N, nV);
    double A[P], Bm[P], Cm[P], Y[P], T1[P], T2[P], Kr[P];
    // one batch of loads: the chunk's rows from the edge table, y and the separators
    const double* E = edge + (size_t)edge_slot(chunk, C) * 4 * P * 64 + v;
#pragma unroll
    for (int t = 0; t < P; t++) {
        if (FULL || t < m) {
            A[t] = E[t * 64];
            Bm[t] = E[(P + t) * 64];
            Cm[t] = E[(2 * P + t) * 64];
            Kr[t] = E[(3 * P + t) * 64];
        }
    }
    const double cp = edge[3 * 4 * P * 64 + edge_slot(chunk, C) * 64 + v];
#pragma unroll
    for (int t = 0; t < P; t++)
        if (FULL || t < m) Y[t] = Yb.ld(v8, s + t);
    WgSep ws;
    if (q_solve >= 0) ws.load(X2, coef, v, G, chunk / 4, kc, q_solve >= 8);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int t = 0; t < P; t++)
        if (FULL || t < m) K[t * 64] = Kr[t];
    double xL = 0.0, xR = 0.0;
    if (q_solve >= 0) ws.get(chunk % 4, kc, xL, xR);
    const double hv = R.h[v];
    int pivots = -1;         // shift class (0: BDF1, 1: BDF2) of the pivots held in T1
    if (q_solve >= 0 && var_active(nsteps, v, q_solve)) {
        if (q_solve >= 8) {
#pragma unroll
            for (int t = 0; t < P; t++)
                if (FULL || t < m) T2[t] = Pb.ld(v8, s + t);
        }
        const BdfCoef bc(q_solve, hv);
        down_pivots<P, FULL>(m, bc.sh, A, Bm, Cm, T1);
        pivots = q_solve >= 8;
        // forward sweep of the right-hand side (T2 = modified rhs)
#pragma unroll
        for (int t = 0; t < P - 1; t++) {
            if (FULL || t < m - 1) {
                double r = bc.rhs(Y[t], q_solve >= 8 ? T2[t] : 0.0, K[t * 64]);
                if (t == 0) r -= A[t] * xL;
                if (t == m - 2) r -= Cm[t] * xR;
                if (t == 0)
                    T2[t] = r;
                else
                    T2[t] = r - A[t] * T1[t - 1] * T2[t - 1];
            }
        }
        // back substitution (yprev takes the old y: the next BDF2 step's y_(n-1))
        const bool save_prev = (q_solve == 0 && R.bdf2) || q_solve >= 8;
        double xn = xR;
#pragma unroll
        for (int t = P - 1; t >= 0; t--) {
            if (FULL || t < m) {
                double xi = xR;
                if (t < m - 1) xi = (t == m - 2) ? T2[t] * T1[t] : (T2[t] - Cm[t] * xn) * T1[t];
                if (save_prev) Pb.st(Y[t], v8, s + t);
                Y[t] = xi;
                Yb.st(xi, v8, s + t);
                xn = xi;
            }
        }
    } else if (q_solve >= 0) {
        return false;  // variable finished its steps: nothing further for it in this stage
    }
    if (q_next < 0 || !var_active(nsteps, v, q_next)) return false;
    if (q_next >= 8) {  // y_(n-1) of the BDF2 step, as stored above or by an earlier solve
#pragma unroll
        for (int t = 0; t < P; t++)
            if (FULL || t < m) T2[t] = Pb.ld(v8, s + t);
    }
    const BdfCoef bc(q_next, hv);
    if (pivots != (q_next >= 8 ? 1 : 0)) down_pivots<P, FULL>(m, bc.sh, A, Bm, Cm, T1);
    // ---- chunk summary for solve q_next: downward sweep of the fill-in column and the
    // right-hand side with the pivots T1, upward elimination of the interior
    double rt = 0.0, at = 0.0;
#pragma unroll
    for (int t = 0; t < P - 1; t++) {
        if (FULL || t < m - 1) {
            const double r = bc.rhs(Y[t], q_next >= 8 ? T2[t] : 0.0, K[t * 64]);
            if (t == 0) {
                rt = r;
                at = A[t];
            } else {
                const double f = A[t] * T1[t - 1];
                rt = r - f * rt;
                at = -f * at;
            }
        }
    }
    double bh = 1.0, rh = 0.0, ch = 0.0, aN = 0.0;
#pragma unroll
    for (int t = P - 2; t >= 0; t--) {
        if (FULL || t < m - 1) {
            const double bs = Bm[t] - bc.sh;
            const double r = bc.rhs(Y[t], q_next >= 8 ? T2[t] : 0.0, K[t * 64]);
            if (t == m - 2) {
                bh = bs;
                rh = r;
                ch = Cm[t];
            } else {
                const double g = Cm[t] * rcp_nr(bh);
                bh = bs - g * aN;
                rh = r - g * rh;
                ch = -g * ch;
            }
            aN = A[t];
        }
    }
    double be = 0.0, re = 0.0, ae = 0.0;
#pragma unroll
    for (int t = 0; t < P; t++) {
        if (t == m - 1) {
            be = Bm[t] - bc.sh;
            re = bc.rhs(Y[t], q_next >= 8 ? T2[t] : 0.0, K[t * 64]);
            ae = A[t];
        }
    }
    double o0 = ae, o1 = be, o2 = re;
    if (m > 1) {
        double ib = 0.0, cL = 0.0;
#pragma unroll
        for (int t = 0; t < P - 1; t++)
            if (t == m - 2) {
                ib = T1[t];
                cL = Cm[t];
            }
        o0 = -ae * at * ib;
        o1 = be - ae * cL * ib;
        o2 = re - ae * rt * ib;
    }
    double n0 = 0.0, n1 = 0.0, n2 = 0.0;
    if (s > 0) {
        n1 = cp;
        if (m > 1) {
            const double ib = rcp_nr(bh);
            n0 = -cp * aN * ib;
            n1 = -cp * ch * ib;
            n2 = -cp * rh * ib;
        }
    }
    sm[0] = o0;
    sm[1] = o1;
    sm[2] = o2;
    sm[3] = n0;
    sm[4] = n1;
    sm[5] = n2;
    return true;
}

// chunk_step for a full chunk away from both ends (rows s-1 .. s+P are generic interior
// rows of DiffusionSystem::get_A), with the same arithmetic as chunk_step<P, true> and one
// memory latency instead of four: every load of the chunk is issued in one batch before
// any is used (the B row index first, since the B row loads depend on it; then the
// splitConst, D and y rows, the separators and, for a BDF2 solve, y_(n-1)), and the
// splitConst rows go to LDS only after the whole batch is in flight.  The downward pivots
// live in LDS (T1[t * 64], written once per shift, read by four passes), which leaves the
// registers to the batch.  A BDF2 solve leaves the old y in registers as the next BDF2
// solve's y_(n-1) instead of reloading it.  BDF2 false: an instance for launches whose
// solves are both BDF1 substeps (q_solve, q_next < 8), without the y_(n-1) registers.
template <int P, bool BDF2>
__device__ __forceinline__ bool chunk_step_interior(const DiffRows& R, const double* __restrict__ g1,
                                                    const double* __restrict__ g2, const int* nsteps,
                                                    int q_solve, int q_next, const double* X2,
                                                    const double* coef, int G, int kc, double* K,
                                                    double* T1, int v, int chunk, double (&sm)[6]) {
    const int N = R.N, nV = R.nV;
    const int s = chunk * P;
    const unsigned v8 = (unsigned)v * sizeof(double);
    const PtBuf Yb(R.y, N, nV), Pb(R.yprev, N, nV), Kb(R.sc, N, nV), Db(R.D, N, nV);
    const bool s2 = BDF2 && q_solve >= 8, n2 = BDF2 && q_next >= 8;
    // ---- the load batch
    const int brow = R.B_row ? R.B_row[v] : v;
    double Kr[P], Dr[P + 2], Y[P], T2[P], Bs[P + 1];
#pragma unroll
    for (int t = 0; t < P; t++) Kr[t] = Kb.ld(v8, s + t);
#pragma unroll
    for (int t = 0; t < P + 2; t++) Dr[t] = Db.ld(v8, s - 1 + t);
#pragma unroll
    for (int t = 0; t < P; t++) Y[t] = Yb.ld(v8, s + t);
    WgSep ws;
    if (q_solve >= 0) ws.load(X2, coef, v, G, chunk / 4, kc, q_solve >= 8);
    if (s2) {
#pragma unroll
        for (int t = 0; t < P; t++) T2[t] = Pb.ld(v8, s + t);
    }
    const double* Brow = R.B + (size_t)brow * N;
#pragma unroll
    for (int t = 0; t < P + 1; t++) Bs[t] = Brow[s - 1 + t];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int t = 0; t < P; t++) K[t * 64] = Kr[t];
    double xL = 0.0, xR = 0.0;
    if (q_solve >= 0) ws.get(chunk % 4, kc, xL, xR);
    // ---- matrix rows (the interior branch of chunk_matrix)
    double A[P], Bm[P], Cm[P];
    double D0 = Dr[1];
    double c2m = (Dr[0] + D0) * g2[s - 1];
    const double cp = Bs[0] * g1[s - 1] * c2m;
#pragma unroll
    for (int t = 0; t < P; t++) {
        const int j = s + t;
        const double Dp = Dr[t + 2];
        const double c1 = Bs[t + 1] * g1[j];
        const double c2 = (D0 + Dp) * g2[j];
        A[t] = c1 * c2m;
        Bm[t] = -c1 * (c2m + c2);
        Cm[t] = c1 * c2;
        c2m = c2;
        D0 = Dp;
    }
    // downward pivots of down_pivots, into LDS
    auto pivots_for = [&](double sh) {
        double tp = 0.0;
#pragma unroll
        for (int t = 0; t < P - 1; t++) {
            const double bs = Bm[t] - sh;
            tp = t == 0 ? rcp_nr(bs) : rcp_nr(bs - A[t] * tp * Cm[t - 1]);
            T1[t * 64] = tp;
        }
    };
    const double hv = R.h[v];
    int pivots = -1;  // shift class (0: BDF1, 1: BDF2) of the pivots held in T1
    if (q_solve >= 0) {
        if (!var_active(nsteps, v, q_solve)) return false;  // variable finished its steps
        const BdfCoef bc(q_solve, hv);
        pivots_for(bc.sh);
        pivots = q_solve >= 8;
#pragma unroll
        for (int t = 0; t < P - 1; t++) {
            double r = bc.rhs(Y[t], s2 ? T2[t] : 0.0, K[t * 64]);
            if (t == 0) r -= A[t] * xL;
            if (t == P - 2) r -= Cm[t] * xR;
            if (t == 0)
                T2[t] = r;
            else
                T2[t] = r - A[t] * T1[(t - 1) * 64] * T2[t - 1];
        }
        const bool save_prev = (q_solve == 0 && R.bdf2) || q_solve >= 8;
        double xn = xR;
#pragma unroll
        for (int t = P - 1; t >= 0; t--) {
            double xi = xR;
            if (t < P - 1)
                xi = (t == P - 2) ? T2[t] * T1[t * 64] : (T2[t] - Cm[t] * xn) * T1[t * 64];
            if (save_prev) Pb.st(Y[t], v8, s + t);
            if (s2) T2[t] = Y[t];  // y_(n-1) of the next BDF2 solve
            Y[t] = xi;
            Yb.st(xi, v8, s + t);
            xn = xi;
        }
    }
    if (q_next < 0 || !var_active(nsteps, v, q_next)) return false;
    if (n2 && q_solve < 8) {  // first BDF2 solve: y_(n-1) as stored by solve 0
#pragma unroll
        for (int t = 0; t < P; t++) T2[t] = Pb.ld(v8, s + t);
    }
    const BdfCoef bc(q_next, hv);
    if (pivots != (q_next >= 8 ? 1 : 0)) pivots_for(bc.sh);
    // ---- chunk summary for solve q_next (as in chunk_step with m = P)
    double rt = 0.0, at = 0.0;
#pragma unroll
    for (int t = 0; t < P - 1; t++) {
        const double r = bc.rhs(Y[t], n2 ? T2[t] : 0.0, K[t * 64]);
        if (t == 0) {
            rt = r;
            at = A[t];
        } else {
            const double f = A[t] * T1[(t - 1) * 64];
            rt = r - f * rt;
            at = -f * at;
        }
    }
    double bh = 1.0, rh = 0.0, ch = 0.0, aN = 0.0;
#pragma unroll
    for (int t = P - 2; t >= 0; t--) {
        const double bs = Bm[t] - bc.sh;
        const double r = bc.rhs(Y[t], n2 ? T2[t] : 0.0, K[t * 64]);
        if (t == P - 2) {
            bh = bs;
            rh = r;
            ch = Cm[t];
        } else {
            const double g = Cm[t] * rcp_nr(bh);
            bh = bs - g * aN;
            rh = r - g * rh;
            ch = -g * ch;
        }
        aN = A[t];
    }
    const double be = Bm[P - 1] - bc.sh;
    const double re = bc.rhs(Y[P - 1], n2 ? T2[P - 1] : 0.0, K[(P - 1) * 64]);
    const double ae = A[P - 1];
    const double ib = T1[(P - 2) * 64];
    const double o0 = -ae * at * ib;
    const double o1 = be - ae * Cm[P - 2] * ib;
    const double o2 = re - ae * rt * ib;
    const double ibh = rcp_nr(bh);
    const double n0 = -cp * aN * ibh;
    const double n1 = -cp * ch * ibh;
    const double n2v = -cp * rh * ibh;
    sm[0] = o0;
    sm[1] = o1;
    sm[2] = o2;
    sm[3] = n0;
    sm[4] = n1;
    sm[5] = n2v;
    return true;
}

// One wave per chunk of P rows [s, e] (e = separator), lane = variable; the kDiffWaves
// chunks of a workgroup form one workgroup chunk (two-level partition, WgSep).
//  q_solve >= 0: solve q_solve's interior rows with the separators from X2 and coef and
//                write y (saving y to yprev first when q_solve starts a BDF step);
//  q_next >= 0:  eliminate the chunks for solve q_next; the chunk summaries meet in LDS and
//                wave 0 eliminates the workgroup's interior separators, writing the
//                workgroup summary (own2, nxt2: plane k, variable v, workgroup g at
//                (k * nV + v) * G + g) and the back-substitution rows (coef).
// Every wave of a workgroup reaches the barrier when q_next >= 0 (q_next is uniform over
// the grid): waves past the last chunk and lanes past nV only skip the work.
template <int P, bool BDF2>
__global__ void __launch_bounds__(256, 2) diff_step_kernel(DiffRows R, const double* __restrict__ g1,
                                                        const double* __restrict__ g2,
                                                        const double* __restrict__ edge,
                                                        const int* nsteps, int q_solve, int q_next,
                                                        const double* X2, double* coef, double* own2,
                                                        double* nxt2) {
    const int v = lane_id();
    // XCD-aware block order: workgroups are dispatched round-robin over the 8 XCDs, so
    // block b runs on XCD b % 8; give the blocks of one XCD consecutive workgroup chunks, so
    // that the summaries that share a cache line (consecutive workgroups of a plane) are
    // written through one L2 and leave it as whole lines
    const int nb = gridDim.x;
    const int g = (nb % 8 == 0) ? (blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8 : blockIdx.x;
    const int w = threadIdx.x / 64;
    const int chunk = __builtin_amdgcn_readfirstlane(g * kDiffWaves + w);
    const int N = R.N, nV = R.nV;
    const int C = (N + P - 1) / P;
    const int G = (C + kDiffWaves - 1) / kDiffWaves;
    const int kc = min(kDiffWaves, C - g * kDiffWaves);  // chunks of this workgroup
    // splitConst rows of the chunk, lane-interleaved in LDS (K[t * 64] = row t of this lane);
    // after the chunk step the region holds the chunk's summary for the workgroup elimination
    __shared__ double kl[kDiffWaves * P * 64];
    double* K = kl + w * P * 64 + v;
    double sm[6];
    bool have = false;
    if (chunk < C && v < nV) {
        if (chunk * P >= 2 && (chunk + 1) * P - 1 <= N - 3) {
            // downward pivots of the interior path, lane-interleaved like K
            __shared__ double tl[kDiffWaves * P * 64];
            double* T1 = tl + w * P * 64 + v;
            have = chunk_step_interior<P, BDF2>(R, g1, g2, nsteps, q_solve, q_next, X2, coef, G, kc, K,
                                                T1, v, chunk, sm);
        } else if ((chunk + 1) * P <= N) {
            have = chunk_step<P, true>(R, edge, nsteps, q_solve, q_next, X2, coef, G, kc, K, v, chunk,
                                       C, sm);
        } else {
            have = chunk_step<P, false>(R, edge, nsteps, q_solve, q_next, X2, coef, G, kc, K, v, chunk,
                                        C, sm);
        }
    }
    if (q_next < 0) return;
    if (have) {
#pragma unroll
        for (int k = 0; k < 6; k++) K[k * 64] = sm[k];
    }
    __syncthreads();
    if (w != 0 || v >= nV || !var_active(nsteps, v, q_next)) return;
    // ---- workgroup elimination: reduced row j of separator u_j (chunk 4g + j),
    // a_j u_(j-1) + b_j u_j + c_j u_(j+1) = r_j with a = own0(j), b = own1(j) + nxt0(j + 1),
    // c = nxt1(j + 1), r = own2(j) + nxt2(j + 1)
    double o[kDiffWaves][3], n[kDiffWaves][3];
#pragma unroll
    for (int j = 0; j < kDiffWaves; j++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            o[j][k] = j < kc ? kl[j * P * 64 + k * 64 + v] : 0.0;
            n[j][k] = j < kc ? kl[j * P * 64 + (3 + k) * 64 + v] : 0.0;
        }
    double O0 = o[0][0], O1 = o[0][1], O2 = o[0][2];
    double N0 = n[0][0], N1 = n[0][1], N2 = n[0][2];
    if (kc == kDiffWaves) {
        // a full workgroup: the same sweeps as below in straight-line code, so that the
        // downward and upward chains of reciprocals can interleave
        double* cr = wg_rhs(coef, g, v);
        const bool first = q_next == 0 || q_next == 8;
        double* cm = wg_mat(coef, G, q_next >= 8, g, v);
        double aj[kWgRows], bj[kWgRows], cj[kWgRows], rj[kWgRows];
#pragma unroll
        for (int j = 0; j < kWgRows; j++) {
            aj[j] = o[j][0];
            bj[j] = o[j][1] + n[j + 1][0];
            cj[j] = n[j + 1][1];
            rj[j] = o[j][2] + n[j + 1][2];
        }
        double bt = bj[0], rt = rj[0], at = aj[0], ib = rcp_nr(bt);
        double bh = bj[kWgRows - 1], rh = rj[kWgRows - 1], ch = cj[kWgRows - 1], aN = aj[kWgRows - 1];
        double rts[kWgRows], ats[kWgRows], ibs[kWgRows];
        rts[0] = rt; ats[0] = at; ibs[0] = ib;
#pragma unroll
        for (int j = 1; j < kWgRows; j++) {
            const double f = aj[j] * ib;
            bt = bj[j] - f * cj[j - 1]; rt = rj[j] - f * rt; at = -f * at;
            ib = rcp_nr(bt);
            rts[j] = rt; ats[j] = at; ibs[j] = ib;
            const int u = kWgRows - 1 - j;
            const double gg = cj[u] * rcp_nr(bh);
            bh = bj[u] - gg * aN; rh = rj[u] - gg * rh; ch = -gg * ch;
            aN = aj[u];
        }
#pragma unroll
        for (int j = 0; j < kWgRows; j++) {
            cr[j * 64] = rts[j];
            if (first) {
                cm[(j * 3 + 0) * 64] = ats[j];
                cm[(j * 3 + 1) * 64] = cj[j];
                cm[(j * 3 + 2) * 64] = ibs[j];
            }
        }
        const double aE = o[kDiffWaves - 1][0], bE = o[kDiffWaves - 1][1], rE = o[kDiffWaves - 1][2];
        O0 = -aE * at * ib;
        O1 = bE - aE * cj[kWgRows - 1] * ib;
        O2 = rE - aE * rt * ib;
        const double cpv = n[0][1], ibh = rcp_nr(bh);
        N0 = n[0][0] - cpv * aN * ibh;
        N1 = -cpv * ch * ibh;
        N2 = n[0][2] - cpv * rh * ibh;
    } else if (kc > 1) {
        double* cr = wg_rhs(coef, g, v);
        // the matrix part, once per shift class (its first solve)
        const bool first = q_next == 0 || q_next == 8;
        double* cm = wg_mat(coef, G, q_next >= 8, g, v);
        // downward over the interior separators j = 0 .. kc-2 (fill-in column at: u_(-1))
        double bt = 0.0, rt = 0.0, at = 0.0, cL = 0.0, ib = 0.0;
#pragma unroll
        for (int j = 0; j < kDiffWaves - 1; j++) {
            if (j < kc - 1) {
                const double aj = o[j][0], bj = o[j][1] + n[j + 1][0], cj = n[j + 1][1],
                             rj = o[j][2] + n[j + 1][2];
                if (j == 0) {
                    bt = bj; rt = rj; at = aj;
                } else {
                    const double f = aj * ib;
                    bt = bj - f * cL; rt = rj - f * rt; at = -f * at;
                }
                cL = cj;
                ib = rcp_nr(bt);
                cr[j * 64] = rt;
                if (first) {
                    cm[(j * 3 + 0) * 64] = at;
                    cm[(j * 3 + 1) * 64] = cj;
                    cm[(j * 3 + 2) * 64] = ib;
                }
            }
        }
        // the workgroup separator's row (own part of its last chunk) without u_(kc-2)
        double aE = 0.0, bE = 0.0, rE = 0.0;
#pragma unroll
        for (int j = 1; j < kDiffWaves; j++)
            if (j == kc - 1) { aE = o[j][0]; bE = o[j][1]; rE = o[j][2]; }
        O0 = -aE * at * ib;
        O1 = bE - aE * cL * ib;
        O2 = rE - aE * rt * ib;
        // upward over the interior separators: the workgroup's part of the row of u_(-1)
        double bh = 0.0, rh = 0.0, ch = 0.0, aN = 0.0;
#pragma unroll
        for (int j = kDiffWaves - 2; j >= 0; j--) {
            if (j < kc - 1) {
                const double aj = o[j][0], bj = o[j][1] + n[j + 1][0], cj = n[j + 1][1],
                             rj = o[j][2] + n[j + 1][2];
                if (j == kc - 2) {
                    bh = bj; rh = rj; ch = cj;
                } else {
                    const double gg = cj * rcp_nr(bh);
                    bh = bj - gg * aN; rh = rj - gg * rh; ch = -gg * ch;
                }
                aN = aj;
            }
        }
        const double cpv = n[0][1], ibh = rcp_nr(bh);
        N0 = n[0][0] - cpv * aN * ibh;
        N1 = -cpv * ch * ibh;
        N2 = n[0][2] - cpv * rh * ibh;
    }
    const size_t pl = (size_t)nV * G, e = (size_t)v * G + g;
    own2[e] = O0;
    own2[pl + e] = O1;
    own2[2 * pl + e] = O2;
    nxt2[e] = N0;
    nxt2[pl + e] = N1;
    nxt2[2 * pl + e] = N2;
}

// Separators of one variable per workgroup of kPcrThreads threads, from the reduced
// tridiagonal system of the workgroup summaries (C rows, C <= kPcrRowsMax).  All threads
// read the rows coalesced into LDS; then each lane of wave 0 owns a run of kRunMax
// consecutive rows (S = ceil(C / kRunMax) <= 64 runs) and eliminates its interior towards
// both ends (the downward/upward elimination of the chunks; in a full run the two sweeps
// are interleaved, independent chains of reciprocals); the S run separators are solved by
// parallel cyclic reduction inside the wave (each level couples row i with rows i -/+ stride
// through cross-lane reads and doubles the stride: no barriers), and each lane
// back-substitutes its run.  The workgroup is the only one of its variable, so its latency
// is the critical path of the solve.
constexpr int kPcrThreads = 512;
constexpr int kRunMax = 16;
constexpr int kPcrRowsMax = kRunMax * 64;
static_assert(kPcrMax / kDiffWaves <= kPcrRowsMax, "reduced system larger than one wave's runs");

// LDS index of reduced row i: one pad double per 16 rows, so that the rows of consecutive
// runs (16 apart) fall into different banks
__host__ __device__ __forceinline__ int pcr_slot(int i) { return i + (i >> 4); }

size_t pcr_lds_doubles(int C) { return 4 * (size_t)pcr_slot(C); }

__global__ void __launch_bounds__(kPcrThreads) pcr_kernel(const double* own, const double* nxt, int C,
                                                          int nV, const int* nsteps, int q,
                                                          double* X) {
    const int v = blockIdx.x;
    if (!var_active(nsteps, v, q)) return;  // uniform over the workgroup
    extern __shared__ double lds[];
    const int CS = pcr_slot(C);
    double* a = lds;
    double* b = a + CS;
    double* c = b + CS;
    double* r = c + CS;
    const size_t pl = (size_t)nV * C;
    const double* ov = own + (size_t)v * C;
    const double* nv = nxt + (size_t)v * C;
    double* Xv = X + (size_t)v * C;
    const int tid = threadIdx.x;
    // reduced row i: own part of workgroup i plus the interior contribution of workgroup
    // i + 1 (all of a thread's loads in flight together; batches past C are skipped)
    constexpr int kLoadBatch = kPcrRowsMax / kPcrThreads;
    {
        double la[kLoadBatch], lb[kLoadBatch], lc[kLoadBatch], lr[kLoadBatch];
#pragma unroll
        for (int k = 0; k < kLoadBatch; k++) {
            if (k * kPcrThreads < C) {
                const int i = min(tid + k * kPcrThreads, C - 1), i1 = min(i + 1, C - 1);
                const double last = i + 1 < C ? 1.0 : 0.0;
                la[k] = ov[i];
                lb[k] = ov[pl + i] + last * nv[i1];
                lc[k] = last * nv[pl + i1];
                lr[k] = ov[2 * pl + i] + last * nv[2 * pl + i1];
            }
        }
#pragma unroll
        for (int k = 0; k < kLoadBatch; k++) {
            const int i = tid + k * kPcrThreads;
            if (k * kPcrThreads < C && i < C) {
                const int o = pcr_slot(i);
                a[o] = la[k];
                b[o] = lb[k];
                c[o] = lc[k];
                r[o] = lr[k];
            }
        }
    }
    __syncthreads();
    if (tid < 64) {
        const int S = (C + kRunMax - 1) / kRunMax;  // runs (separator rows), <= 64
        const bool act = tid < S;
        const int s0 = tid * kRunMax;
        const int e0 = act ? min(s0 + kRunMax, C) - 1 : s0;  // separator of this run
        const int mi = e0 - s0;                               // interior rows of the run
        // the run's rows in registers (rows past the run are clamped copies, never used)
        double ra[kRunMax], rb[kRunMax], rc[kRunMax], rr[kRunMax];
#pragma unroll
        for (int k = 0; k < kRunMax; k++) {
            const int o = pcr_slot(min(s0 + k, C - 1));
            ra[k] = a[o];
            rb[k] = b[o];
            rc[k] = c[o];
            rr[k] = r[o];
        }
        const double cp = tid > 0 && act ? c[pcr_slot(s0 - 1)] : 0.0;
        const bool full = mi == kRunMax - 1;
        double iv[kRunMax - 1];  // full runs: reciprocals of the downward pivots
        double aE = 0.0, bE = 0.0, rE = 0.0;
#pragma unroll
        for (int k = 0; k < kRunMax; k++)
            if (k == mi) {
                aE = ra[k];
                bE = rb[k];
                rE = rr[k];
            }
        // downward over the interior: the separator row in terms of the previous separator;
        // upward over the interior: the run's contribution to the previous separator's row
        double bt = 1.0, rt = 0.0, at = 0.0, cL = 0.0;
        double bh = 1.0, rh = 0.0, ch = 0.0, aN = 0.0;
        if (full) {
            // the downward pivots' reciprocals stay in iv for the back-substitution, whose
            // forward sweep has the same pivots
            constexpr int M = kRunMax - 1;
            bt = rb[0]; rt = rr[0]; at = ra[0]; cL = rc[0];
            bh = rb[M - 1]; rh = rr[M - 1]; ch = rc[M - 1]; aN = ra[M - 1];
#pragma unroll
            for (int k = 1; k < M; k++) {
                iv[k - 1] = rcp_nr(bt);
                const double f = ra[k] * iv[k - 1];
                bt = rb[k] - f * cL; rt = rr[k] - f * rt; at = -f * at;
                cL = rc[k];
                const int u = M - 1 - k;
                const double g = rc[u] * rcp_nr(bh);
                bh = rb[u] - g * aN; rh = rr[u] - g * rh; ch = -g * ch;
                aN = ra[u];
            }
            iv[M - 1] = rcp_nr(bt);
        } else {
#pragma unroll
            for (int k = 0; k < kRunMax; k++) {
                if (k < mi) {
                    if (k == 0) {
                        bt = rb[k]; rt = rr[k]; at = ra[k];
                    } else {
                        const double f = ra[k] * rcp_nr(bt);
                        bt = rb[k] - f * cL; rt = rr[k] - f * rt; at = -f * at;
                    }
                    cL = rc[k];
                }
            }
#pragma unroll
            for (int k = kRunMax - 1; k >= 0; k--) {
                if (k < mi) {
                    if (k == mi - 1) {
                        bh = rb[k]; rh = rr[k]; ch = rc[k];
                    } else {
                        const double g = rc[k] * rcp_nr(bh);
                        bh = rb[k] - g * aN; rh = rr[k] - g * rh; ch = -g * ch;
                    }
                    aN = ra[k];
                }
            }
        }
        // separator row of the run: (sa, sb, sc, sr); runs past S are identity rows
        double sa = aE, sb = bE, sc = 0.0, sr = rE;
        if (mi > 0) {
            const double ib = full ? iv[kRunMax - 2] : rcp_nr(bt);
            sa = -aE * at * ib;
            sb = bE - aE * cL * ib;
            sr = rE - aE * rt * ib;
        }
        double nb_ = 0.0, nc_ = cp, nr_ = 0.0;
        if (tid > 0 && mi > 0) {
            const double ib = rcp_nr(bh);
            nb_ = -cp * aN * ib;
            nc_ = -cp * ch * ib;
            nr_ = -cp * rh * ib;
        }
        if (!act) {
            sa = 0.0; sb = 1.0; sr = 0.0;
            nb_ = 0.0; nc_ = 0.0; nr_ = 0.0;
        }
        // the next run's contribution to this separator's row
        {
            const double nbn = __shfl_down(nb_, 1), ncn = __shfl_down(nc_, 1), nrn = __shfl_down(nr_, 1);
            if (tid + 1 < S) {
                sb += nbn;
                sc = ncn;
                sr += nrn;
            }
        }
        // parallel cyclic reduction on the S separator rows, one row per lane
        for (int st = 1; st < S; st <<= 1) {
            const int lo = tid - st, hi = tid + st;
            const int l = lo >= 0 ? lo : tid, h = hi < 64 ? hi : tid;
            const double la_ = __shfl(sa, l), lb_ = __shfl(sb, l), lc_ = __shfl(sc, l), lr_ = __shfl(sr, l);
            const double ha_ = __shfl(sa, h), hb_ = __shfl(sb, h), hc_ = __shfl(sc, h), hr_ = __shfl(sr, h);
            double na = 0.0, nc = 0.0, nb = sb, nr = sr;
            if (lo >= 0) {
                const double al = -sa * rcp_nr(lb_);
                na = al * la_;
                nb += al * lc_;
                nr += al * lr_;
            }
            if (hi < S) {
                const double ga = -sc * rcp_nr(hb_);
                nc = ga * hc_;
                nb += ga * ha_;
                nr += ga * hr_;
            }
            sa = na;
            sb = nb;
            sc = nc;
            sr = nr;
        }
        // back-substitute each run's interior between its neighbouring separators
        const double xR = sr / sb;
        const double xLn = __shfl_up(xR, 1);
        const double xL = tid > 0 ? xLn : 0.0;
        if (act) {
            if (full) {
                // the pivots of the forward elimination are the downward sweep's (iv)
                constexpr int M = kRunMax - 1;
#pragma unroll
                for (int k = 0; k < M; k++) {
                    double ri = rr[k];
                    if (k == 0) ri -= ra[k] * xL;
                    if (k == M - 1) ri -= rc[k] * xR;
                    if (k > 0) {
                        const double f = ra[k] * iv[k - 1];
                        ri -= f * rr[k - 1];
                    }
                    rr[k] = ri;
                }
                double xn = xR;
#pragma unroll
                for (int k = M - 1; k >= 0; k--) {
                    const double xi = (k == M - 1) ? rr[k] * iv[k] : (rr[k] - rc[k] * xn) * iv[k];
                    r[pcr_slot(s0 + k)] = xi;
                    xn = xi;
                }
            } else {
                // forward elimination with 1/diag (in ra) and swept right-hand sides (in rr),
                // then back substitution
#pragma unroll
                for (int k = 0; k < kRunMax; k++) {
                    if (k < mi) {
                        double ri = rr[k];
                        if (k == 0) ri -= ra[k] * xL;
                        if (k == mi - 1) ri -= rc[k] * xR;
                        if (k == 0) {
                            ra[k] = rcp_nr(rb[k]);
                        } else {
                            const double f = ra[k] * ra[k - 1];
                            ra[k] = rcp_nr(rb[k] - f * rc[k - 1]);
                            ri -= f * rr[k - 1];
                        }
                        rr[k] = ri;
                    }
                }
                double xn = xR;
#pragma unroll
                for (int k = kRunMax - 1; k >= 0; k--) {
                    if (k < mi) {
                        const double xi = (k == mi - 1) ? rr[k] * ra[k] : (rr[k] - rc[k] * xn) * ra[k];
                        r[pcr_slot(s0 + k)] = xi;
                        xn = xi;
                    }
                }
            }
            r[pcr_slot(e0)] = xR;
        }
    }
    __syncthreads();
    // coalesced store of the separators (each run wrote its rows' values into the r slots)
    for (int i = tid; i < C; i += kPcrThreads) Xv[i] = r[pcr_slot(i)];
}

constexpr int P1 = 16;   // level-1 chunk (points)
constexpr int P2 = 32;   // level-2 chunk (separators)
constexpr int kThomasMax = 512;

template <class Rows>
int run_solve(const Rows& R, StateOut out, int n, int nV, const int* nsteps, int q,
              DiffWork& w, hipStream_t st) {
    const int wpb = 4;
    const int C1 = (n + P1 - 1) / P1;
    hipLaunchKernelGGL((summary_kernel<P1, Rows>), dim3((C1 + wpb - 1) / wpb), dim3(64 * wpb), 0, st,
                       R, n, nV, nsteps, q, w.own1, w.nxt1);
    ReducedRows R2{w.own1, w.nxt1, C1, nV};
    SepOut sep1{w.X1, nV};
    if (C1 <= kThomasMax) {
        hipLaunchKernelGGL((thomas_kernel<ReducedRows, SepOut>), dim3((nV + 63) / 64), dim3(64), 0, st,
                           R2, sep1, C1, nV, nsteps, q, w.cp, w.dp);
    } else {
        const int C2 = (C1 + P2 - 1) / P2;
        hipLaunchKernelGGL((summary_kernel<P2, ReducedRows>), dim3((C2 + wpb - 1) / wpb), dim3(64 * wpb),
                           0, st, R2, C1, nV, nsteps, q, w.own2, w.nxt2);
        ReducedRows R3{w.own2, w.nxt2, C2, nV};
        SepOut sep2{w.X2, nV};
        hipLaunchKernelGGL((thomas_kernel<ReducedRows, SepOut>), dim3((nV + 63) / 64), dim3(64), 0, st,
                           R3, sep2, C2, nV, nsteps, q, w.cp, w.dp);
        hipLaunchKernelGGL((solve_kernel<P2, ReducedRows, SepOut>), dim3((C2 + wpb - 1) / wpb),
                           dim3(64 * wpb), 0, st, R2, sep1, C1, nV, nsteps, q, w.X2);
    }
    hipLaunchKernelGGL((solve_kernel<P1, Rows, StateOut>), dim3((C1 + wpb - 1) / wpb), dim3(64 * wpb),
                       0, st, R, out, n, nV, nsteps, q, w.X1);
    return (int)hipGetLastError();
}

}  // namespace

size_t diffusion_work_doubles(int nV, int N) {
    const int C1 = (N + P1 - 1) / P1;
    const int C2 = (C1 + P2 - 1) / P2;
    // own1, nxt1 (3 each), X1, own2, nxt2, X2, cp, dp, hmin chunk mins, yprev; grid metrics;
    // edge table; workgroup back-substitution rows of the fused path
    return (size_t)nV * (6 * C1 + C1 + 6 * C2 + C2 + 2 * (size_t)std::max(C1, C2) + C1 + N) +
           2 * (size_t)N + edge_doubles<P1>() + wg_coef_doubles((C1 + kDiffWaves - 1) / kDiffWaves);
}

void diffusion_carve(double* base, int nV, int N, DiffWork& w) {
    const int C1 = (N + P1 - 1) / P1;
    const int C2 = (C1 + P2 - 1) / P2;
    const size_t Cm = (size_t)std::max(C1, C2);
    double* p = base;
    w.own1 = p; p += (size_t)3 * C1 * nV;
    w.nxt1 = p; p += (size_t)3 * C1 * nV;
    w.X1 = p; p += (size_t)C1 * nV;
    w.own2 = p; p += (size_t)3 * C2 * nV;
    w.nxt2 = p; p += (size_t)3 * C2 * nV;
    w.X2 = p; p += (size_t)C2 * nV;
    w.cp = p; p += Cm * nV;
    w.dp = p; p += Cm * nV;
    w.cmin = p; p += (size_t)C1 * nV;
    w.yprev = p; p += (size_t)N * nV;
    w.g1 = p; p += N;
    w.g2 = p; p += N;
    w.edge = p; p += edge_doubles<P1>();
    w.coef = p; p += wg_coef_doubles((C1 + kDiffWaves - 1) / kDiffWaves);
}

namespace {
// per-variable step sizes and step counts, delivered as kernel arguments so that the
// values are fixed at launch (no host buffer has to outlive the call)
__global__ void set_steps_kernel(DiffSteps s, int nV, double* h, int* ns) {
    const int v = threadIdx.x;
    if (v < nV) {
        h[v] = s.h[v];
        ns[v] = s.ns[v];
    }
}
}  // namespace

int launch_set_steps(const DiffSteps& s, int nV, double* h, int* ns, hipStream_t st) {
    hipLaunchKernelGGL(set_steps_kernel, dim3(1), dim3(64), 0, st, s, nV, h, ns);
    return (int)hipGetLastError();
}

int launch_diffusion_hmin(const DiffArgs& a, DiffWork& w, double* hmin, hipStream_t st) {
    DiffRows R{a.B, a.B_row, a.D, a.sc, a.x, nullptr, nullptr, nullptr, a.N, a.nV, a.alpha,
               a.left, a.right, a.yInf, a.wallConst, a.wallVar, 0};
    const int wpb = 4;
    const int C1 = (a.N + P1 - 1) / P1;
    hipLaunchKernelGGL(hmin_chunk_kernel, dim3((C1 + wpb - 1) / wpb), dim3(64 * wpb), 0, st, R, P1,
                       w.cmin);
    hipLaunchKernelGGL(hmin_reduce_kernel, dim3((a.nV + 63) / 64), dim3(64), 0, st, w.cmin, C1, a.nV,
                       hmin);
    return (int)hipGetLastError();
}

int launch_diffusion_solve(const DiffArgs& a, double* y, const double* h, const int* nsteps, int q,
                           DiffWork& w, hipStream_t st) {
    DiffRows R{a.B, a.B_row, a.D, a.sc, a.x, y, w.yprev, h, a.N, a.nV, a.alpha,
               a.left, a.right, a.yInf, a.wallConst, a.wallVar, q};
    StateOut out{y, w.yprev, a.nV, q == 0 || q >= 8};
    return run_solve(R, out, a.N, a.nV, nsteps, q, w, st);
}

int launch_diffusion_stage(const DiffArgs& a, double* y, const DiffSteps& steps, double* h, int* nsteps,
                           int nsolve, DiffWork& w, hipStream_t st) {
    const int C1 = (a.N + P1 - 1) / P1;
    // the fused path addresses point-major arrays with 32-bit buffer offsets
    const bool fits32 = (size_t)a.N * a.nV * sizeof(double) < ((size_t)1 << 31);
    if (C1 > kPcrMax || a.N < 2 * P1 || !fits32) {
        if (int e = launch_set_steps(steps, a.nV, h, nsteps, st)) return e;
        for (int q = 0; q < nsolve; q++) {
            int e = launch_diffusion_solve(a, y, h, nsteps, q, w, st);
            if (e) return e;
        }
        return 0;
    }
    DiffRows R{a.B, a.B_row, a.D, a.sc, a.x, y, w.yprev, h, a.N, a.nV, a.alpha,
               a.left, a.right, a.yInf, a.wallConst, a.wallVar, 0};
    // y_0 is kept in yprev only for the stage's first BDF2 solve (solve 8)
    R.bdf2 = nsolve > 8;
    const int wpb = kDiffWaves;
    const int G = (C1 + wpb - 1) / wpb;  // workgroup chunks: rows of the reduced system
    const dim3 grid(G), block(64 * wpb);
    const size_t pcr_lds = pcr_lds_doubles(G) * sizeof(double);
    const int nbm = (a.N + 255) / 256;
    hipLaunchKernelGGL((stage_setup_kernel<P1>), dim3(nbm + (3 * (P1 + 1) * 64 + 255) / 256), dim3(256), 0,
                       st, R, nbm, w.g1, w.g2, w.edge, steps, h, nsteps);
    // the BDF1 instance for launches whose solves are both BDF1 substeps
    auto step = [&](int qs, int qn) {
        if (qs >= 8 || qn >= 8)
            hipLaunchKernelGGL((diff_step_kernel<P1, true>), grid, block, 0, st, R, w.g1, w.g2, w.edge,
                               nsteps, qs, qn, w.X1, w.coef, w.own1, w.nxt1);
        else
            hipLaunchKernelGGL((diff_step_kernel<P1, false>), grid, block, 0, st, R, w.g1, w.g2, w.edge,
                               nsteps, qs, qn, w.X1, w.coef, w.own1, w.nxt1);
    };
    step(-1, 0);
    for (int q = 0; q < nsolve; q++) {
        hipLaunchKernelGGL(pcr_kernel, dim3(a.nV), dim3(kPcrThreads), pcr_lds, st, w.own1, w.nxt1, G, a.nV,
                           nsteps, q, w.X1);
        step(q, q + 1 < nsolve ? q + 1 : -1);
    }
    return (int)hipGetLastError();
}

int launch_explicit_solve(const double* A, const double* B, const double* Cc, const double* K,
                          int nSys, int N, double* y, const double* h, const int* nsteps, int q,
                          DiffWork& w, hipStream_t st) {
    ExplicitRows R{A, B, Cc, K, y, w.yprev, h, N, nSys, q};
    StateOut out{y, w.yprev, nSys, q == 0 || q >= 8};
    return run_solve(R, out, N, nSys, nsteps, q, w, st);
}

}  // namespace emx
