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
// up, and finally every chunk solves its interior against known separators.
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
    return q < 8 || (q - 7) < nsteps[v];
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
    // own1, nxt1 (3 each), X1, own2, nxt2, X2, cp, dp, hmin chunk mins, yprev
    return (size_t)nV * (6 * C1 + C1 + 6 * C2 + C2 + 2 * (size_t)std::max(C1, C2) + C1 + N);
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

int launch_explicit_solve(const double* A, const double* B, const double* Cc, const double* K,
                          int nSys, int N, double* y, const double* h, const int* nsteps, int q,
                          DiffWork& w, hipStream_t st) {
    ExplicitRows R{A, B, Cc, K, y, w.yprev, h, N, nSys, q};
    StateOut out{y, w.yprev, nSys, q == 0 || q >= 8};
    return run_solve(R, out, N, nSys, nsteps, q, w, st);
}

}  // namespace emx
