// Plain GPU port of the chemistry stage, used as the comparison baseline: one thread
// per grid point running the serial QssIntegrator + SourceSystemQSS loop, with all
// per-point arrays in a global workspace laid out [array][point] so that neighbouring
// threads touch neighbouring addresses.  Same arithmetic as the wavefront kernel;
// no LDS, no cooperation between lanes, one launch over all points.
#include <hip/hip_runtime.h>

#include "emx_common.h"
#include "kernels.h"

namespace emx {

namespace {

struct Ws {
    double* base;
    int n, p;
    __device__ __forceinline__ double& operator()(int slot) const { return base[(size_t)slot * n + p]; }
};

// workspace slots
struct Slots {
    int y, q, d, rtau, rtaus, y1, ys, qs, scratch, C, mu, hk, kf, rkc, ropf, ropr, Mg, wQ, wD, klow,
        khigh, work, total;
    __host__ __device__ Slots(int nv, int K, int S, int G) {
        int o = 0;
        y = o; o += nv; q = o; o += nv; d = o; o += nv; rtau = o; o += nv; rtaus = o; o += nv;
        y1 = o; o += nv; ys = o; o += nv; qs = o; o += nv; scratch = o; o += nv;
        C = o; o += K; mu = o; o += K; hk = o; o += K;
        kf = o; o += S; rkc = o; o += S; ropf = o; o += S; ropr = o; o += S;
        Mg = o; o += (G > 0 ? G : 1);
        wQ = o; o += K; wD = o; o += K;
        klow = o; o += S; khigh = o; o += S; work = o; o += S;
        total = o;
    }
};

__device__ double prodC(uint32_t st, const Ws& w, int C0) {
    int n = st_n(st);
    if (n == 1) return w(C0 + st_k(st, 0));
    if (n == 2) return w(C0 + st_k(st, 0)) * w(C0 + st_k(st, 1));
    return w(C0 + st_k(st, 0)) * w(C0 + st_k(st, 1)) * w(C0 + st_k(st, 2));
}

struct PlainState {
    double T, rho, cp, qdot;
};

// SourceSystemQSS::odefun for one point, serial
__device__ void plain_odefun(const DevMech& m, const Ws& w, const Slots& s, const double* perturb,
                             const double* sc, double qloss, bool corrector, PlainState& P) {
    const int K = m.K, S = m.nRounds * 64;
    if (!corrector) P.T = w(s.y + 1);
    const double T = P.T;
    double norm = 0.0;
    for (int k = 0; k < K; k++) norm += fmax(w(s.y + 2 + k), 0.0);
    double sum = 0.0;
    for (int k = 0; k < K; k++) {
        double yk = fmax(w(s.y + 2 + k), 0.0) * (1.0 / norm);
        double ym = yk * m.rW[k];
        w(s.C + k) = ym;
        sum += ym;
    }
    const double mmw = 1.0 / sum;
    const double dens = m.P * mmw / (T * kGasConstant);
    for (int k = 0; k < K; k++) w(s.C + k) = w(s.C + k) * dens;
    const double ctot = dens / mmw;
    if (!corrector) {
        const double logT = log(T), recipT = 1.0 / T, RT = T * kGasConstant;
        const double pressure = kGasConstant * (dens / mmw) * T;
        const double tmp = log(pressure / m.Pref) * RT;
        double scp = 0.0;
        for (int k = 0; k < K; k++) {
            const double* c = m.nasa + 14 * k + ((T <= m.Tmid[k]) ? 0 : 7);
            double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
            double ct0 = c[0], ct1 = c[1] * T, ct2 = c[2] * T2, ct3 = c[3] * T3, ct4 = c[4] * T4;
            double cpR = ct0 + ct1 + ct2 + ct3 + ct4;
            double hRT = ct0 + 0.5 * ct1 + 1.0 / 3.0 * ct2 + 0.25 * ct3 + 0.2 * ct4 + c[5] * (1.0 / T);
            double sR = ct0 * logT + ct1 + 0.5 * ct2 + 1.0 / 3.0 * ct3 + 0.25 * ct4 + c[6];
            w(s.mu + k) = RT * (hRT - sR) + tmp;
            w(s.hk + k) = RT * hRT;
            scp += (w(s.C + k) / dens) * cpR;
        }
        P.rho = dens;
        P.cp = (kGasConstant * (mmw * scp)) / mmw;
        const double logStandConc = log(pressure / RT), rrt = 1.0 / RT;
        for (int i = 0; i < S; i++) {
            int info = m.info[i], kind = info_kind(info);
            double kf = 0.0, rkc = 0.0;
            if (kind == kNone) { w(s.kf + i) = 0.0; w(s.rkc + i) = 0.0; continue; }
            if (info_rev(info)) {
                uint32_t pr = m.prod[i], re = m.reac[i];
                double dg = 0.0;
                for (int j = 0; j < st_n(pr); j++) dg += w(s.mu + st_k(pr, j));
                for (int j = 0; j < st_n(re); j++) dg -= w(s.mu + st_k(re, j));
                rkc = fmin(exp(dg * rrt - m.dn[i] * logStandConc), kBigNumber);
            }
            if (kind <= kThreeBody) kf = m.A[i] * exp(m.b[i] * logT - m.EaR[i] * recipT);
            if (kind == kLindemann || kind == kTroe) {
                w(s.klow + i) = m.A0[i] * exp(m.b0[i] * logT - m.EaR0[i] * recipT);
                w(s.khigh + i) = m.A[i] * exp(m.b[i] * logT - m.EaR[i] * recipT);
                double work = 0.0;
                if (kind == kTroe) {
                    double Fc = (1.0 - m.troeA[i]) * exp(-T * m.rT3[i]) + m.troeA[i] * exp(-T * m.rT1[i]);
                    if (m.T2[i] != 0.0) Fc += exp(-m.T2[i] / T);
                    work = log10(fmax(Fc, kSmallNumber));
                }
                w(s.work + i) = work;
            }
            w(s.kf + i) = kf;
            w(s.rkc + i) = rkc;
        }
    }
    for (int g = 0; g < m.nGroups; g++) {
        double sum2 = 0.0;
        for (int e = m.grp_off[g]; e < m.grp_off[g + 1]; e++) sum2 += m.grp_eff[e] * w(s.C + m.grp_sp[e]);
        w(s.Mg + g) = m.grp_def[g] * ctot + sum2;
    }
    for (int i = 0; i < S; i++) {
        int info = m.info[i], kind = info_kind(info);
        double ropf = 0.0, ropr = 0.0;
        if (kind != kNone) {
            double kf = w(s.kf + i);
            if (kind == kLindemann || kind == kTroe) {
                double khigh = w(s.khigh + i);
                double pr = w(s.Mg + info_group(info)) * w(s.klow + i) / (khigh + kSmallNumber);
                double F = 1.0;
                if (kind == kTroe) {
                    double work = w(s.work + i);
                    double lpr = log10(fmax(pr, kSmallNumber));
                    double cc = -0.4 - 0.67 * work, nn = 0.75 - 1.27 * work;
                    double f1 = (lpr + cc) / (nn - 0.14 * (lpr + cc));
                    F = pow(10.0, work / (1.0 + f1 * f1));
                }
                pr *= F / (1.0 + pr);
                kf = pr * khigh;
            }
            ropf = kf * perturb[i];
            if (kind == kThreeBody) ropf *= w(s.Mg + info_group(info));
            ropr = ropf;
            ropf *= prodC(m.reac[i], w, s.C);
            ropr *= w(s.rkc + i);
            if (info_rev(info)) ropr *= prodC(m.prod[i], w, s.C);
        }
        w(s.ropf + i) = ropf;
        w(s.ropr + i) = ropr;
    }
    double qd = 0.0;
    for (int k = 0; k < K; k++) {
        double q = 0.0, d = 0.0;
        for (int c = m.sp_chunk0[k]; c < m.sp_chunk0[k] + m.sp_nchunk[k]; c++) {
            for (int e = c * kChunkLen; e < (c + 1) * kChunkLen; e++) {
                uint32_t en = m.entry[e];
                if (en_nu(en) == 0) continue;  // chunk padding
                double nu = (double)en_nu(en), f = w(s.ropf + en_slot(en)), r = w(s.ropr + en_slot(en));
                if (en_prod(en)) { q += nu * f; d += nu * r; } else { q += nu * r; d += nu * f; }
            }
        }
        w(s.wQ + k) = q;
        w(s.wD + k) = d;
        qd += (q - d) * w(s.hk + k);
    }
    P.qdot = -qd;
    w(s.q + 0) = sc ? sc[0] : 0.0;
    w(s.d + 0) = 0.0;
    w(s.q + 1) = P.qdot / (P.rho * P.cp) + (sc ? sc[1] : 0.0);
    w(s.d + 1) = qloss / (P.rho * P.cp);
    for (int k = 0; k < K; k++) {
        w(s.q + 2 + k) = w(s.wQ + k) * m.W[k] / P.rho + (sc ? sc[2 + k] : 0.0);
        w(s.d + 2 + k) = w(s.wD + k) * m.W[k] / P.rho;
    }
}

__global__ void __launch_bounds__(64) qss_plain_kernel(DevMech m, DevOpts o, const double* state_in,
                                                       const double* split_const, const double* qloss,
                                                       const double* perturb, double tf, int nPoints,
                                                       double* state_out, int* status, int* stat1,
                                                       int* stat2, double* wsbase) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nPoints) return;
    const int nv = m.nVars;
    Slots s(nv, m.K, m.nRounds * 64, m.nGroups);
    Ws w{wsbase, nPoints, p};
    const double* sc = split_const ? split_const + (size_t)p * nv : nullptr;
    const double ql = qloss ? qloss[p] : 0.0;
    PlainState P{0, 0, 0, 0};
    for (int i = 0; i < nv; i++) {
        double v = state_in[(size_t)p * nv + i];
        w(s.y + i) = (i == 0) ? v : fmax(v, o.qss_minval);
    }
    auto ymin = [&](int i) { return i == 0 ? -1e4 : o.qss_minval; };
    const double tfd = 1.000008;
    double tn = 0, ts = 0, dt = 0;
    bool first = true;
    int gcount = 0, rcount = 0, err = 0;
    long budget = 20000000L;
    while (tfd * tn < tf && err == 0) {
        plain_odefun(m, w, s, perturb, sc, ql, false, P);
        gcount++;
        if (first) {
            first = false;
            double scrtch = 1e-25;
            for (int i = 0; i < nv; i++) {
                double y = w(s.y + i), q = w(s.q + i), d = w(s.d + i);
                if (fabs(y) > o.qss_abstol) {
                    double absq = fabs(q);
                    double sg = (0.1 * o.qss_epsmin * absq - d) > 0 ? 1.0 : (0.1 * o.qss_epsmin * absq - d) < 0 ? -1.0 : 0.0;
                    double scr2 = fabs(1 / y) * sg;
                    scrtch = fmax(fmax(scr2 * d, -fabs(absq - d) * scr2), scrtch);
                }
            }
            dt = fmin(fmin(0.5 / scrtch, tf), o.qss_dtmax);
        }
        ts = tn;
        for (int i = 0; i < nv; i++) {
            w(s.rtau + i) = i ? dt * w(s.d + i) / w(s.y + i) : 0.0;
            w(s.qs + i) = w(s.q + i);
            w(s.ys + i) = w(s.y + i);
            w(s.rtaus + i) = w(s.rtau + i);
        }
        while (true) {
            if (--budget < 0) { err = -3; break; }
            for (int i = 0; i < nv; i++) {
                double r = w(s.rtau + i);
                w(s.scratch + i) = (w(s.q + i) - w(s.d + i)) /
                                   (1.0 + r * (180 + r * (60 + r * (11 + r))) / (360 + r * (60 + r * (12 + r))));
            }
            for (int iter = 0; iter < o.qss_itermax; iter++) {
                for (int i = 0; i < nv; i++) w(s.y + i) = fmax(ymin(i), w(s.ys + i) + dt * w(s.scratch + i));
                if (iter == 0) {
                    tn = ts + dt;
                    for (int i = 0; i < nv; i++) w(s.y1 + i) = w(s.y + i);
                }
                plain_odefun(m, w, s, perturb, sc, ql, true, P);
                gcount++;
                for (int i = 0; i < nv; i++) {
                    double rtaub = i ? 0.5 * (w(s.rtaus + i) + dt * w(s.d + i) / w(s.y + i)) : 0.0;
                    double a = (180. + rtaub * (60. + rtaub * (11. + rtaub))) / (360. + rtaub * (60. + rtaub * (12. + rtaub)));
                    w(s.scratch + i) = (w(s.qs + i) * (1.0 - a) + w(s.q + i) * a - w(s.ys + i) * rtaub / dt) / (1.0 + a * rtaub);
                }
            }
            double eps = 1e-10;
            for (int i = 0; i < nv; i++) {
                double scr2 = w(s.ys + i) + dt * w(s.scratch + i);
                if (i) scr2 = fmax(scr2, ymin(i));
                double scr1 = fabs(scr2 - w(s.y1 + i));
                double y = fmax(scr2, ymin(i));
                w(s.y + i) = y;
                if (fabs(y) > o.qss_abstol && 0.25 * (w(s.ys + i) + y) > ymin(i)) {
                    scr1 = scr1 / y;
                    double q = w(s.q + i), d = w(s.d + i);
                    eps = fmax(.5 * (scr1 + fmin(fabs(q - d) / (q + d + 1e-30), scr1)), eps);
                }
            }
            eps /= o.qss_epsmin;
            if (dt <= o.qss_dtmin + 1e-16 * tn) { err = -1; break; }
            if (!(eps == eps)) { err = -4; break; }
            bool ok = eps <= o.qss_epsmax;
            if (ok && tf <= tn * tfd) break;
            if (!ok) tn = ts;
            double rteps = 0.5 * (eps + 1.0);
            rteps = 0.5 * (rteps + eps / rteps);
            rteps = 0.5 * (rteps + eps / rteps);
            double dto = dt;
            dt = fmin(fmin(dt * (1.0 / rteps + 0.005), tfd * (tf - tn)), o.qss_dtmax);
            if (!ok) {
                rcount++;
                for (int i = 0; i < nv; i++) w(s.rtaus + i) *= dt / dto;
            } else {
                break;
            }
        }
    }
    for (int i = 0; i < nv; i++)
        state_out[(size_t)p * nv + i] = err >= 0 ? w(s.y + i) : state_in[(size_t)p * nv + i];
    if (status) status[p] = err;
    if (stat1) stat1[p] = gcount;
    if (stat2) stat2[p] = rcount;
}

}  // namespace

size_t plain_workspace_doubles(const DevMech& m, int nPoints) {
    Slots s(m.nVars, m.K, m.nRounds * 64, m.nGroups);
    return (size_t)s.total * nPoints;
}

int launch_qss_plain(const DevMech& m, const DevOpts& o, const double* state_in,
                     const double* split_const, const double* qloss, const double* perturb,
                     double dt, int nPoints, double* state_out, int* status, int* stat1, int* stat2,
                     double* ws, hipStream_t stream) {
    hipLaunchKernelGGL(qss_plain_kernel, dim3((nPoints + 63) / 64), dim3(64), 0, stream, m, o, state_in,
                       split_const, qloss, perturb, dt, nPoints, state_out, status, stat1, stat2, ws);
    return (int)hipGetLastError();
}

}  // namespace emx
