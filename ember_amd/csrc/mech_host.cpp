// Mechanism table reader and device layout builder.
#include "mech_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>

namespace emx {

std::string read_table(const std::string& path, HostMech& m) {
    FILE* f = fopen(path.c_str(), "r");
    if (!f) return "cannot open mechanism table " + path;
    char tag[64], buf[256];
    int E = 0;
    std::string err;
    if (fscanf(f, "%63s %255s", tag, buf) != 2 || strcmp(tag, "EMXM") != 0) err = "not an EMXM table";
    while (err.empty() && fscanf(f, "%63s", tag) == 1) {
        if (!strcmp(tag, "END")) break;
        if (!strcmp(tag, "name")) {
            if (fscanf(f, "%255s", buf) != 1) err = "bad name";
            m.name = buf;
        } else if (!strcmp(tag, "counts")) {
            if (fscanf(f, "%d %d %d", &m.K, &m.I, &E) != 3) err = "bad counts";
        } else if (!strcmp(tag, "pref")) {
            if (fscanf(f, "%lf", &m.Pref) != 1) err = "bad pref";
        } else if (!strcmp(tag, "elements")) {
            for (int e = 0; e < E && err.empty(); e++)
                if (fscanf(f, "%255s", buf) != 1) err = "bad elements";
        } else if (!strcmp(tag, "S")) {
            double v[18];
            if (fscanf(f, "%255s", buf) != 1) { err = "bad species"; break; }
            for (int j = 0; j < 18 && err.empty(); j++)
                if (fscanf(f, "%lf", &v[j]) != 1) err = "bad species data";
            for (int e = 0; e < E && err.empty(); e++) {
                double c;
                if (fscanf(f, "%lf", &c) != 1) err = "bad composition";
            }
            m.species.push_back(buf);
            m.W.push_back(v[0]);
            m.Tmid.push_back(v[2]);
            for (int j = 0; j < 14; j++) m.nasa.push_back(v[4 + j]);
        } else if (!strcmp(tag, "R")) {
            HostRxn r;
            int neff = 0;
            if (fscanf(f, "%d %d %d %d %d %d %d %d %d %d", &r.type, &r.rev, &r.nr, &r.r[0], &r.r[1],
                       &r.r[2], &r.np, &r.p[0], &r.p[1], &r.p[2]) != 10 ||
                fscanf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %d %lf %d", &r.A, &r.b, &r.EaR,
                       &r.A0, &r.b0, &r.EaR0, &r.troeA, &r.T3, &r.T1, &r.T2, &r.hasT2, &r.defeff,
                       &neff) != 13) {
                err = "bad reaction";
                break;
            }
            for (int j = 0; j < neff; j++) {
                int k;
                double e;
                if (fscanf(f, "%d %lf", &k, &e) != 2) { err = "bad efficiencies"; break; }
                r.eff.emplace_back(k, e);
            }
            if (r.type < 0 || r.type > 6) err = "unknown reaction type";
            m.rxn.push_back(r);
        } else if (!strcmp(tag, "PLOG")) {
            // the PLOG rates of the preceding reaction
            int n = 0;
            if (m.rxn.empty() || m.rxn.back().type != 4 || fscanf(f, "%d", &n) != 1 || n < 1) {
                err = "bad PLOG record";
                break;
            }
            std::vector<double>& v = m.rxn.back().plog;
            v.resize(4 * (size_t)n);
            for (int j = 0; j < 4 * n && err.empty(); j++)
                if (fscanf(f, "%lf", &v[j]) != 1) err = "bad PLOG data";
            for (int j = 1; j < n && err.empty(); j++)
                if (v[4 * j] < v[4 * (j - 1)]) err = "PLOG pressures must not decrease";
        } else if (!strcmp(tag, "ORD")) {
            // explicit forward orders of the preceding reaction (irreversible)
            HostRxn* r = m.rxn.empty() ? nullptr : &m.rxn.back();
            if (!r || r->rev || fscanf(f, "%lf %lf %lf", &r->ord[0], &r->ord[1], &r->ord[2]) != 3 ||
                r->ord[0] < 0 || r->ord[1] < 0 || r->ord[2] < 0) {
                err = "bad ORD record";
                break;
            }
            r->hasOrd = 1;
        } else if (!strcmp(tag, "SRI")) {
            HostRxn* r = m.rxn.empty() ? nullptr : &m.rxn.back();
            if (!r || r->type != 6 ||
                fscanf(f, "%lf %lf %lf %lf %lf", &r->sri[0], &r->sri[1], &r->sri[2], &r->sri[3], &r->sri[4]) != 5) {
                err = "bad SRI record";
                break;
            }
        } else if (!strcmp(tag, "CHEB")) {
            HostRxn* r = m.rxn.empty() ? nullptr : &m.rxn.back();
            if (!r || r->type != 5 ||
                fscanf(f, "%d %d %lf %lf %lf %lf", &r->chebT, &r->chebP, &r->chebTmin, &r->chebTmax,
                       &r->chebPmin, &r->chebPmax) != 6 || r->chebT < 1 || r->chebP < 1) {
                err = "bad CHEB record";
                break;
            }
            r->cheb.resize((size_t)r->chebT * r->chebP);
            for (auto& c : r->cheb)
                if (err.empty() && fscanf(f, "%lf", &c) != 1) err = "bad CHEB data";
        } else if (!strcmp(tag, "TRANSPORT")) {
            if (fscanf(f, "%lf %lf", &m.trTmin, &m.trTmax) != 2 || m.K <= 0) {
                err = "bad TRANSPORT record";
                break;
            }
            m.visc.assign((size_t)m.K * 5, 0.0);
            m.cond.assign((size_t)m.K * 5, 0.0);
            m.diff.assign((size_t)m.K * (m.K + 1) / 2 * 5, 0.0);
        } else if (!strcmp(tag, "TV") || !strcmp(tag, "TC") || !strcmp(tag, "TD")) {
            int j = 0, k = 0;
            const bool pair = tag[1] == 'D';
            if (!m.has_transport() || fscanf(f, "%d", &j) != 1 || (pair && fscanf(f, "%d", &k) != 1) ||
                j < 0 || j >= m.K || k < 0 || k >= m.K || (pair && k < j)) {
                err = "bad transport fit record";
                break;
            }
            double* c = tag[1] == 'V' ? &m.visc[(size_t)j * 5]
                      : tag[1] == 'C' ? &m.cond[(size_t)j * 5] : &m.diff[(size_t)pair_index(j, k, m.K) * 5];
            for (int i = 0; i < 5 && err.empty(); i++)
                if (fscanf(f, "%lf", &c[i]) != 1) err = "bad transport fit data";
        } else {
            err = std::string("unknown record ") + tag;
        }
    }
    fclose(f);
    if (err.empty() && ((int)m.species.size() != m.K || (int)m.rxn.size() != m.I))
        err = "table counts do not match records";
    return err;
}

static uint32_t pack_stoich(int n, const int* k) {
    uint32_t s = (uint32_t)n;
    for (int j = 0; j < 3; j++) s |= ((uint32_t)(j < n ? k[j] : kPadSpecies) & 0xff) << (8 * (j + 1));
    return s;
}

std::string build_layout(const HostMech& m, DevLayout& L) {
    if (m.K + 2 > kMaxVars) return "mechanism has more than 62 species";
    if (m.K > 255) return "too many species";
    // device order: Troe, Lindemann, three-body, elementary (stable)
    std::vector<int> order(m.I);
    for (int i = 0; i < m.I; i++) order[i] = i;
    auto rank = [&](int i) {
        int t = m.rxn[i].type;
        return t == 3 ? 0 : (t == 2 || t == 6) ? 1 : t == 1 ? 2 : 3;  // PLOG, Chebyshev with elementary
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return rank(a) < rank(b); });
    int nFall = 0;
    for (int i = 0; i < m.I; i++) nFall += (m.rxn[i].type == 2 || m.rxn[i].type == 3 || m.rxn[i].type == 6);
    L.nRounds = (m.I + 63) / 64;
    L.nFallRounds = (nFall + 63) / 64;
    if (L.nRounds > kMaxRounds) return "mechanism has more than 512 reactions";
    if (L.nFallRounds > kMaxFallRounds) return "mechanism has more than 128 falloff reactions";
    const int S = L.nRounds * 64;
    L.orig.assign(S, -1);
    L.slot_of_rxn.assign(m.I, -1);
    for (int s = 0; s < m.I; s++) {
        L.orig[s] = order[s];
        L.slot_of_rxn[order[s]] = s;
    }
    L.A.assign(S, 0); L.b.assign(S, 0); L.EaR.assign(S, 0); L.dn.assign(S, 0);
    L.A0.assign(S, 0); L.b0.assign(S, 0); L.EaR0.assign(S, 0);
    L.troeA.assign(S, 0); L.rT3.assign(S, 0); L.rT1.assign(S, 0); L.T2.assign(S, 0);
    L.reac.assign(S, 0); L.prod.assign(S, 0); L.info.assign(S, kNone);
    L.ord.assign(3 * (size_t)S, 0.0);
    // third-body efficiency groups
    std::map<std::pair<double, std::vector<std::pair<int, double>>>, int> groups;
    std::vector<std::pair<double, std::vector<std::pair<int, double>>>> glist;
    for (int s = 0; s < m.I; s++) {
        const HostRxn& r = m.rxn[order[s]];
        L.A[s] = r.A; L.b[s] = r.b; L.EaR[s] = r.EaR;
        L.dn[s] = double(r.np) - double(r.nr);
        L.A0[s] = r.A0; L.b0[s] = r.b0; L.EaR0[s] = r.EaR0;
        L.troeA[s] = r.troeA;
        // TroeRate: 1/T3, 1/T1 (an zero T3/T1 means an infinite rate constant)
        L.rT3[s] = (std::fabs(r.T3) < 1e-300) ? INFINITY : 1.0 / r.T3;
        L.rT1[s] = (std::fabs(r.T1) < 1e-300) ? INFINITY : 1.0 / r.T1;
        L.T2[s] = r.T2;
        L.reac[s] = pack_stoich(r.nr, r.r);
        L.prod[s] = pack_stoich(r.np, r.p);
        int group = 255;
        if ((r.type >= 1 && r.type <= 3) || r.type == 6) {
            std::vector<std::pair<int, double>> e;
            for (auto& pe : r.eff) e.emplace_back(pe.first, pe.second - r.defeff);
            auto key = std::make_pair(r.defeff, e);
            auto it = groups.find(key);
            if (it == groups.end()) {
                group = (int)glist.size();
                groups[key] = group;
                glist.push_back(key);
            } else {
                group = it->second;
            }
        }
        int kind = r.type == 3 ? kTroe : r.type == 2 ? kLindemann : r.type == 1 ? kThreeBody
                 : r.type == 4 ? kPlog : r.type == 5 ? kCheb : r.type == 6 ? kSri : kElementary;
        L.info[s] = kind | (r.rev ? 8 : 0) | (r.hasT2 ? 16 : 0) | (r.hasOrd ? kOrdBit : 0) | (group << 8);
        if (r.hasOrd)
            for (int j = 0; j < 3; j++) L.ord[(size_t)j * S + s] = j < r.nr ? r.ord[j] : 0.0;
    }
    if ((int)glist.size() > 63) return "more than 63 distinct third-body efficiency sets";
    L.nGroups = (int)glist.size();
    L.grp_off.assign(1, 0);
    for (auto& g : glist) {
        L.grp_def.push_back(g.first);
        for (auto& e : g.second) {
            L.grp_sp.push_back(e.first);
            L.grp_eff.push_back(e.second);
        }
        L.grp_off.push_back((int)L.grp_sp.size());
    }
    if (L.grp_sp.empty()) { L.grp_sp.push_back(0); L.grp_eff.push_back(0.0); }
    if (L.grp_def.empty()) L.grp_def.push_back(0.0);
    // species-sum entries: products then reactants, slot order.  Indices address the
    // rates of progress as one array (forward rates, zero slot, reverse rates), so that
    // a product entry adds ropf to creation and ropr to destruction, a reactant entry the
    // other way round
    const uint32_t S1 = (uint32_t)(L.nRounds * 64 + 1);
    auto make_entry = [&](uint32_t slot, uint32_t nu, bool product) {
        const uint32_t f = slot, r = S1 + slot;
        return en_pack(product ? f : r, product ? r : f, nu);
    };
    std::vector<std::vector<uint32_t>> ent(m.K);
    for (int s = 0; s < m.I; s++) {
        const HostRxn& r = m.rxn[order[s]];
        std::map<int, int> cnt;
        for (int j = 0; j < r.np; j++) cnt[r.p[j]]++;
        for (auto& c : cnt) ent[c.first].push_back(make_entry((uint32_t)s, (uint32_t)c.second, true));
    }
    for (int s = 0; s < m.I; s++) {
        const HostRxn& r = m.rxn[order[s]];
        std::map<int, int> cnt;
        for (int j = 0; j < r.nr; j++) cnt[r.r[j]]++;
        for (auto& c : cnt) ent[c.first].push_back(make_entry((uint32_t)s, (uint32_t)c.second, false));
    }
    // fixed-length chunks of kChunkLen entries; a species' chunks are consecutive
    const uint32_t zs = (uint32_t)(L.nRounds * 64);
    const uint32_t zero_entry = en_pack(zs, zs, 0);  // zero slot, multiplicity 0
    L.sp_chunk0.assign(m.K, 0);
    L.sp_nchunk.assign(m.K, 0);
    L.entry.clear();
    int nch = 0;
    for (int k = 0; k < m.K; k++) {
        L.sp_chunk0[k] = nch;
        size_t n = ent[k].size();
        int nc = std::max<int>(1, (int)((n + kChunkLen - 1) / kChunkLen));
        for (int c = 0; c < nc; c++)
            for (int e = 0; e < kChunkLen; e++) {
                size_t i = (size_t)c * kChunkLen + e;
                L.entry.push_back(i < n ? ent[k][i] : zero_entry);
            }
        L.sp_nchunk[k] = nc;
        nch += nc;
    }
    L.nChunks = nch;
    L.nChunkRounds = (nch + 63) / 64;
    if (L.nChunkRounds > kMaxChunkRounds) return "species-sum chunking does not fit";
    L.entry.resize((size_t)L.nChunkRounds * 64 * kChunkLen, zero_entry);
    L.chunk_sp.assign((size_t)L.nChunkRounds * 64, 0);
    for (int k = 0; k < m.K; k++)
        for (int c = 0; c < L.sp_nchunk[k]; c++) L.chunk_sp[L.sp_chunk0[k] + c] = k;
    // adjoint tables
    L.grp_slot_off.assign(1, 0);
    L.grp_slot.clear();
    for (int g = 0; g < L.nGroups; g++) {
        for (int s = 0; s < m.I; s++)
            if (info_group(L.info[s]) == g) L.grp_slot.push_back(s);
        L.grp_slot_off.push_back((int)L.grp_slot.size());
    }
    if (L.grp_slot.empty()) L.grp_slot.push_back(0);
    L.sp_grp_off.assign(1, 0);
    L.sp_grp.clear();
    L.sp_grp_eff.clear();
    for (int k = 0; k < m.K; k++) {
        for (int g = 0; g < L.nGroups; g++)
            for (int e = L.grp_off[g]; e < L.grp_off[g + 1]; e++)
                if (L.grp_sp[e] == k) {
                    L.sp_grp.push_back(g);
                    L.sp_grp_eff.push_back(L.grp_eff[e]);
                }
        L.sp_grp_off.push_back((int)L.sp_grp.size());
    }
    if (L.sp_grp.empty()) { L.sp_grp.push_back(0); L.sp_grp_eff.push_back(0.0); }
    L.rW.resize(m.K);
    for (int k = 0; k < m.K; k++) L.rW[k] = 1.0 / m.W[k];
    return "";
}

static void host_nasa(const HostMech& m, int k, double T, double& cpR, double& hRT, double& sR) {
    const double* c = &m.nasa[14 * k + ((T <= m.Tmid[k]) ? 0 : 7)];
    double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
    double ct0 = c[0], ct1 = c[1] * T, ct2 = c[2] * T2, ct3 = c[3] * T3, ct4 = c[4] * T4;
    cpR = ct0 + ct1 + ct2 + ct3 + ct4;
    hRT = ct0 + 0.5 * ct1 + 1.0 / 3.0 * ct2 + 0.25 * ct3 + 0.2 * ct4 + c[5] * (1.0 / T);
    sR = ct0 * std::log(T) + ct1 + 0.5 * ct2 + 1.0 / 3.0 * ct3 + 0.25 * ct4 + c[6];
}

void build_pressure_tables(const HostMech& m, const DevLayout& L, double P, std::vector<int32_t>& aux,
                           std::vector<double>& auxd) {
    const int S = L.nRounds * 64;
    aux.assign(S, -1);
    auxd.clear();
    const double logP = std::log(P), log10P = std::log10(P);
    for (int s = 0; s < m.I; s++) {
        const HostRxn& r = m.rxn[L.orig[s]];
        if (r.type == 4) {
            // PlogRate at fixed P: the levels (distinct pressures) bracketing P, the end
            // level on both sides outside the listed range
            const int n = (int)r.plog.size() / 4;
            std::vector<int> first;  // index of the first rate of each level
            for (int j = 0; j < n; j++)
                if (j == 0 || r.plog[4 * j] != r.plog[4 * (j - 1)]) first.push_back(j);
            const int nl = (int)first.size();
            auto level_end = [&](int l) { return l + 1 < nl ? first[l + 1] : n; };
            int lo = 0, hi = 0;
            double wP = 0.0;
            if (logP >= std::log(r.plog[4 * first[nl - 1]])) {
                lo = hi = nl - 1;
            } else if (logP >= std::log(r.plog[0])) {
                hi = 0;
                while (std::log(r.plog[4 * first[hi]]) <= logP) hi++;  // upper_bound
                lo = hi - 1;
                const double logP1 = std::log(r.plog[4 * first[lo]]), logP2 = std::log(r.plog[4 * first[hi]]);
                const double rDeltaP = 1.0 / (logP2 - logP1);
                wP = (logP - logP1) * rDeltaP;
            }
            aux[s] = (int32_t)auxd.size();
            const int n1 = level_end(lo) - first[lo], n2 = level_end(hi) - first[hi];
            auxd.push_back(wP);
            auxd.push_back(n1);
            auxd.push_back(n2);
            for (int l : {lo, hi})
                for (int j = first[l]; j < level_end(l); j++) {
                    const double A = r.plog[4 * j + 1];
                    auxd.push_back(A);
                    auxd.push_back(A > 0 ? std::log(A) : -INFINITY);
                    auxd.push_back(r.plog[4 * j + 2]);
                    auxd.push_back(r.plog[4 * j + 3]);
                }
        } else if (r.type == 5) {
            // ChebyshevRate at fixed P: the pressure sums d_t = sum_p a_tp T_p(Pr)
            const double PrNum = -std::log10(r.chebPmin) - std::log10(r.chebPmax);
            const double PrDen = 1.0 / (std::log10(r.chebPmax) - std::log10(r.chebPmin));
            const double Pr = (2 * log10P + PrNum) * PrDen;
            std::vector<double> d(r.chebT, 0.0);
            double Cnm1 = Pr, Cn = 1.0;
            for (int p = 0; p < r.chebP; p++) {
                const double Cnp1 = 2 * Pr * Cn - Cnm1;
                for (int t = 0; t < r.chebT; t++) d[t] += Cn * r.cheb[(size_t)t * r.chebP + p];
                Cnm1 = Cn;
                Cn = Cnp1;
            }
            aux[s] = (int32_t)auxd.size();
            auxd.push_back(r.chebT);
            auxd.push_back(-1.0 / r.chebTmin - 1.0 / r.chebTmax);
            auxd.push_back(1.0 / (1.0 / r.chebTmax - 1.0 / r.chebTmin));
            auxd.insert(auxd.end(), d.begin(), d.end());
        } else if (r.type == 6) {
            aux[s] = (int32_t)auxd.size();
            auxd.insert(auxd.end(), r.sri, r.sri + 5);
        }
    }
    if (auxd.empty()) auxd.push_back(0.0);
}

void build_interp_tables(const HostMech& m, const DevLayout& L, double P, int nT, double Tmin,
                         double Tmax, std::vector<double>& kf_c, std::vector<double>& kf_s,
                         std::vector<double>& kc_c, std::vector<double>& kc_s) {
    const int S = L.nRounds * 64;
    kf_c.assign((size_t)nT * S, 0.0);
    kf_s.assign((size_t)nT * S, 0.0);
    kc_c.assign((size_t)nT * S, 0.0);
    kc_s.assign((size_t)nT * S, 0.0);
    std::vector<double> mu(m.K);
    for (int n = 0; n < nT; n++) {
        double T = Tmin + ((double)n) / (nT - 1) * (Tmax - Tmin);
        double RT = T * kGasConstant;
        double tmp = std::log(P / m.Pref) * RT;
        for (int k = 0; k < m.K; k++) {
            double cpR, hRT, sR;
            host_nasa(m, k, T, cpR, hRT, sR);
            mu[k] = RT * (hRT - sR) + tmp;
        }
        double logStandConc = std::log(P / RT), rrt = 1.0 / RT;
        double logT = std::log(T), recipT = 1.0 / T;
        for (int s = 0; s < m.I; s++) {
            const HostRxn& r = m.rxn[L.orig[s]];
            double rkc = 0.0;
            if (r.rev) {
                double dg = 0.0;
                if (r.np == 1) dg += mu[r.p[0]];
                else if (r.np == 2) dg += mu[r.p[0]] + mu[r.p[1]];
                else dg += mu[r.p[0]] + mu[r.p[1]] + mu[r.p[2]];
                if (r.nr == 1) dg -= mu[r.r[0]];
                else if (r.nr == 2) dg -= (mu[r.r[0]] + mu[r.r[1]]);
                else dg -= (mu[r.r[0]] + mu[r.r[1]] + mu[r.r[2]]);
                rkc = std::min(std::exp(dg * rrt - L.dn[s] * logStandConc), kBigNumber);
            }
            kc_c[(size_t)n * S + s] = rkc;
            if (r.type <= 1) kf_c[(size_t)n * S + s] = r.A * std::exp(r.b * logT - r.EaR * recipT);
        }
    }
    // the last node keeps slope 0, as InterpKinetics leaves it (src/chemistry0d.cpp:327-331)
    double dT = (Tmax - Tmin) / (nT - 1);
    for (int n = 0; n < nT - 1; n++)
        for (int s = 0; s < S; s++) {
            size_t a = (size_t)n * S + s, b = (size_t)(n + 1) * S + s;
            kf_s[a] = (kf_c[b] - kf_c[a]) / dT;
            kc_s[a] = (kc_c[b] - kc_c[a]) / dT;
        }
}

}  // namespace emx
