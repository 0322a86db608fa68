// PENNANT hydro oracle -- TEST INFRASTRUCTURE ONLY.
//
// A serial CPU restatement of pennant-legion's LOC_PROC task bodies, run as
// one piece (`-n 1`).  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only as the reference the GPU
// results are compared with (or as the timed CPU baseline); the product path
// never links or calls it.
//
// Build: `make -C oracle` (g++ -O2 -ffp-contract=off, no -march=native), which
// produces oracle/liboracle_pennant.so.  Every routine cites the reference
// file:line whose arithmetic it restates; expression trees are kept in the
// reference's evaluation order so that results are reproducible bit-for-bit.
//
// Pinning (what checks each part of this oracle against something other than
// its own formulas) is recorded in tests/test_oracle_*.py:
//   * whole rect cycle      -> reference golden test/sedov/sedov.xy (+ .geo t)
//   * geometry              -> closed forms on undeformed rect and pie meshes
//   * EOS / PolyGas / TTS   -> closed forms and zero-net-force invariants
//   * QCS                   -> zero-viscosity special cases, textbook zkur
//   * corner sums, BCs      -> mass conservation, projection identities
//   * dt logic              -> brute-force scalar cases
//   * OMP variant           -> agreement with the serial variant
// Pie-mesh triangle branches and the radial-velocity init are pinned only by
// closed forms and invariants (no reference-held fixture covers them).

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <sstream>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// double2 value semantics: src/Vec2.hh:24-203.  Division of a vector by a
// scalar multiplies by the reciprocal (src/Vec2.hh:145-149).
struct V2 {
    double x, y;
};
inline V2 mk(double x, double y) { V2 v; v.x = x; v.y = y; return v; }
inline V2 operator+(V2 a, V2 b) { return mk(a.x + b.x, a.y + b.y); }
inline V2 operator-(V2 a, V2 b) { return mk(a.x - b.x, a.y - b.y); }
inline V2 operator-(V2 a) { return mk(-a.x, -a.y); }
inline V2 operator*(V2 v, double r) { return mk(v.x * r, v.y * r); }
inline V2 operator*(double r, V2 v) { return mk(v.x * r, v.y * r); }
inline V2 operator/(V2 v, double r) { double rinv = 1. / r; return mk(v.x * rinv, v.y * rinv); }
inline double dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
inline double cross(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
inline double length(V2 v) { return std::sqrt(v.x * v.x + v.y * v.y); }
inline double length2(V2 v) { return v.x * v.x + v.y * v.y; }
inline V2 rotateCCW(V2 v) { return mk(-v.y, v.x); }
inline V2 project(V2 v, V2 u) { return v - dot(v, u) * u; }

// std::min / std::max as the reference calls them (libstdc++ definitions).
inline double smin(double a, double b) { return (b < a) ? b : a; }
inline double smax(double a, double b) { return (a < b) ? b : a; }

// ---------------------------------------------------------------------------
// .pnt key/value deck: src/InputFile.cc:24-104.
struct Deck {
    std::map<std::string, std::string> kv;
    std::string err;

    bool parse(const char* text) {
        std::istringstream in(text);
        std::string line;
        while (true) {
            // getline + eof test as src/InputFile.cc:35-38: a last line with
            // no terminating newline is not read
            std::getline(in, line);
            if (in.eof()) break;
            std::istringstream ls(line);
            std::string key;
            ls >> key;
            if (key.empty() || key[0] == '#') continue;
            if (kv.count(key)) { err = "Duplicate key " + key + " in input file"; return false; }
            std::string val;
            std::getline(ls, val);
            kv[key] = val;
        }
        return true;
    }
    bool has(const std::string& k) const { return kv.count(k) != 0; }
    double getd(const std::string& k, double d) const {
        auto it = kv.find(k); if (it == kv.end()) return d;
        std::istringstream s(it->second); double v = 0.; s >> v; return v;
    }
    int geti(const std::string& k, int d) const {
        auto it = kv.find(k); if (it == kv.end()) return d;
        std::istringstream s(it->second); int v = 0; s >> v; return v;
    }
    std::string gets(const std::string& k, const std::string& d) const {
        auto it = kv.find(k); if (it == kv.end()) return d;
        std::istringstream s(it->second); std::string v; s >> v; return v;
    }
    std::vector<double> getl(const std::string& k) const {
        std::vector<double> out;
        auto it = kv.find(k); if (it == kv.end()) return out;
        std::istringstream s(it->second); double v; while (s >> v) out.push_back(v);
        return out;
    }
};

struct BC {
    V2 vfix;
    std::vector<int64_t> pts;
};

struct Ora {
    // parameters (src/Driver.cc:78-87, src/Hydro.cc:206-214, src/PolyGas.cc:58-59,
    // src/TTS.cc:46-47, src/QCS.cc:84-86, src/Mesh.cc:218-219)
    int cstop; double tstop, dtmax, dtinit, dtfac; int dtreport;
    double cfl, cflv, rinit, einit, rinitsub, einitsub, uinitradial;
    std::vector<double> bcx, bcy, subregion;
    double pg_gamma, pg_ssmin, tts_alfa, tts_ssmin, qgamma, q1, q2;
    std::string meshtype; int64_t nzx, nzy; double lenx, leny;

    int64_t nump = 0, numz = 0, nums = 0;
    std::vector<int64_t> mapsp1, mapsp2, mapsz, mapss3, mapss4;
    std::vector<int> znump;

    std::vector<V2> px, pxp, px0, pu, pu0, pf, pap;
    std::vector<double> pmaswt;
    std::vector<V2> zx, zxp, zuc;
    std::vector<double> zarea, zvol, zareap, zvolp, zvol0, zdl, zm, zr, zrp, ze,
        zetot, zw, zwrate, zp, zss, zdu, ztmp;
    std::vector<V2> ex, exp_, ssurfp, sfp, sfq, sft, cqe1, cqe2;
    std::vector<double> sarea, svol, sareap, svolp, elen, smf, carea, cevol, cdu,
        cdiv, ccos, crmu, cw;
    std::vector<BC> bcs;

    // driver state
    double time = 0., dt = 0., dtrec = 0.;   // dtrec is the f_cdt future
    int cycle = 0;
    int numsbad_pred = 0, numsbad_corr = 0;
    double dtnew = 0., dvovmax = 0.;
    bool use_omp = false;
    std::string err;
};

// ---------------------------------------------------------------------------
// Mesh generation, one piece.  Numbering and coordinates follow
// src/GenMesh.cc:469-526 (counts), 790-873 (points), 898-975 (zones),
// 987-1282 (sides) with numpcx = numpcy = 1, where point compaction is the
// identity (src/Mesh.cc:1504-1528).

bool gen_mesh(Ora& o) {
    const int64_t nzx = o.nzx, nzy = o.nzy, npx = nzx + 1;
    if (o.meshtype == "rect") {
        o.nump = (nzx + 1) * (nzy + 1);
        o.numz = nzx * nzy;
        o.nums = 4 * o.numz;
    } else if (o.meshtype == "pie") {
        o.nump = (nzx + 1) * nzy + 1;
        o.numz = nzx * nzy;
        o.nums = 4 * nzx * (nzy - 1) + 3 * nzx;
    } else {
        o.err = "meshtype " + o.meshtype + " is not supported (hex is assert(false) in the reference)";
        return false;
    }
    o.px.resize(o.nump);
    o.znump.resize(o.numz);
    o.mapsp1.resize(o.nums); o.mapsp2.resize(o.nums); o.mapsz.resize(o.nums);
    o.mapss3.resize(o.nums); o.mapss4.resize(o.nums);

    if (o.meshtype == "rect") {
        // src/GenMesh.cc:790-825
        const double dx = o.lenx / (double)nzx;
        const double dy = o.leny / (double)nzy;
        for (int64_t p = 0; p < o.nump; ++p) {
            const int64_t i = p % npx, j = p / npx;
            o.px[p] = mk(dx * double(i), dy * double(j));
        }
        // src/GenMesh.cc:573-580 (ZNUMP fill = 4)
        for (int64_t z = 0; z < o.numz; ++z) o.znump[z] = 4;
        // src/GenMesh.cc:987-1098
        for (int64_t s = 0; s < o.nums; ++s) {
            const int64_t side = s % 4;
            o.mapss3[s] = (side == 0) ? s + 3 : s - 1;
            o.mapss4[s] = (side == 3) ? s - 3 : s + 1;
            const int64_t zone = s / 4;
            o.mapsz[s] = zone;
            const int64_t zidx = zone % nzx, zidy = zone / nzx;
            int64_t x1 = 0, x2 = 0, y1 = 0, y2 = 0;
            switch (side) {
                case 0: x1 = zidx;     x2 = zidx + 1; y1 = zidy;     y2 = zidy;     break;
                case 1: x1 = zidx + 1; x2 = zidx + 1; y1 = zidy;     y2 = zidy + 1; break;
                case 2: x1 = zidx + 1; x2 = zidx;     y1 = zidy + 1; y2 = zidy + 1; break;
                default: x1 = zidx;    x2 = zidx;     y1 = zidy + 1; y2 = zidy;     break;
            }
            o.mapsp1[s] = y1 * npx + x1;
            o.mapsp2[s] = y2 * npx + x2;
        }
    } else {
        // src/GenMesh.cc:842-873 (pie points; x = theta, y = r)
        const double dth = o.lenx / (double)nzx;
        const double dr = o.leny / (double)nzy;
        o.px[0] = mk(0., 0.);
        for (int64_t p = 1; p < o.nump; ++p) {
            const int64_t i = (p - 1) % npx, j = (p - 1) / npx + 1;
            const double th = dth * (double)(nzx - i);
            const double r = dr * (double)j;
            // src/GenMesh.cc:866-867 evaluates cos(th) and sin(th); the
            // reference's g++ -O2 build (DEBUG=0, Makefile:29) fuses the pair
            // into one glibc sincos call, whose results differ from separate
            // sin/cos calls in the last bit for some angles.
            double sn, cs;
            sincos(th, &sn, &cs);
            o.px[p] = mk(r * cs, r * sn);
        }
        // src/GenMesh.cc:936-975
        for (int64_t z = 0; z < o.numz; ++z) o.znump[z] = (z / nzx == 0) ? 3 : 4;
        // src/GenMesh.cc:1100-1282
        for (int64_t s = 0; s < o.nums; ++s) {
            int64_t zone, side;
            if (s < 3 * nzx) { zone = s / 3; side = s % 3; }
            else { const int64_t t = s - 3 * nzx; zone = nzx + t / 4; side = t % 4; }
            o.mapsz[s] = zone;
            const int64_t zidx = zone % nzx, zidy = zone / nzx;
            if (zidy == 0) {
                o.mapss3[s] = (side == 0) ? s + 2 : s - 1;
                o.mapss4[s] = (side == 2) ? s - 2 : s + 1;
                const int64_t p0 = zidx + 1;
                switch (side) {
                    case 0: o.mapsp1[s] = 0;      o.mapsp2[s] = p0 + 1; break;
                    case 1: o.mapsp1[s] = p0 + 1; o.mapsp2[s] = p0;     break;
                    default: o.mapsp1[s] = p0;    o.mapsp2[s] = 0;      break;
                }
            } else {
                o.mapss3[s] = (side == 0) ? s + 3 : s - 1;
                o.mapss4[s] = (side == 3) ? s - 3 : s + 1;
                const int64_t p0 = (zidy - 1) * npx + zidx + 1;
                switch (side) {
                    case 0: o.mapsp1[s] = p0;           o.mapsp2[s] = p0 + 1;       break;
                    case 1: o.mapsp1[s] = p0 + 1;       o.mapsp2[s] = p0 + npx + 1; break;
                    case 2: o.mapsp1[s] = p0 + npx + 1; o.mapsp2[s] = p0 + npx;     break;
                    default: o.mapsp1[s] = p0 + npx;    o.mapsp2[s] = p0;           break;
                }
            }
        }
    }
    return true;
}

// ---------------------------------------------------------------------------
// Geometry tasks (src/Mesh.cc).

// calcCtrsTask, src/Mesh.cc:797-837
void calc_ctrs(Ora& o, const std::vector<V2>& px, std::vector<V2>& ex, std::vector<V2>& zx) {
    for (int64_t z = 0; z < o.numz; ++z) zx[z] = mk(0., 0.);
    for (int64_t s = 0; s < o.nums; ++s) {
        const V2 px1 = px[o.mapsp1[s]];
        const V2 px2 = px[o.mapsp2[s]];
        ex[s] = 0.5 * (px1 + px2);
        const int64_t z = o.mapsz[s];
        const int n = o.znump[z];
        zx[z] = zx[z] + px1 / (double)n;
    }
}

// calcVolsTask, src/Mesh.cc:928-989 (returns the bad-side count)
int calc_vols(Ora& o, const std::vector<V2>& px, const std::vector<V2>& zx,
              std::vector<double>& sa_out, std::vector<double>& sv_out,
              std::vector<double>& za, std::vector<double>& zv) {
    for (int64_t z = 0; z < o.numz; ++z) { za[z] = 0.; zv[z] = 0.; }
    const double third = 1. / 3.;
    int count = 0;
    for (int64_t s = 0; s < o.nums; ++s) {
        const V2 px1 = px[o.mapsp1[s]];
        const V2 px2 = px[o.mapsp2[s]];
        const int64_t z = o.mapsz[s];
        const V2 c = zx[z];
        const double sa = 0.5 * cross(px2 - px1, c - px1);
        const double sv = third * sa * (px1.x + px2.x + c.x);
        sa_out[s] = sa;
        sv_out[s] = sv;
        za[z] += sa;
        zv[z] += sv;
        if (sv <= 0.) count += 1;
    }
    return count;
}

// calcSideFracsTask, src/Mesh.cc:1085-1101
void calc_side_fracs(Ora& o) {
    for (int64_t s = 0; s < o.nums; ++s) o.smf[s] = o.sarea[s] / o.zarea[o.mapsz[s]];
}

// calcSurfVecsTask, src/Mesh.cc:1104-1123
void calc_surf_vecs(Ora& o) {
    for (int64_t s = 0; s < o.nums; ++s)
        o.ssurfp[s] = rotateCCW(o.exp_[s] - o.zxp[o.mapsz[s]]);
}

// calcEdgeLenTask, src/Mesh.cc:1151-1179
void calc_edge_len(Ora& o) {
    for (int64_t s = 0; s < o.nums; ++s)
        o.elen[s] = length(o.pxp[o.mapsp2[s]] - o.pxp[o.mapsp1[s]]);
}

// calcCharLenTask, src/Mesh.cc:1216-1244
void calc_char_len(Ora& o) {
    for (int64_t z = 0; z < o.numz; ++z) o.zdl[z] = 1.e99;
    for (int64_t s = 0; s < o.nums; ++s) {
        const int64_t z = o.mapsz[s];
        const double area = o.sareap[s];
        const double base = o.elen[s];
        const int np = o.znump[z];
        const double fac = (np == 3 ? 3. : 4.);
        const double sdl = fac * area / base;
        o.zdl[z] = smin(o.zdl[z], sdl);
    }
}

// ---------------------------------------------------------------------------
// Hydro tasks (src/Hydro.cc).

// advPosHalfTask, src/Hydro.cc:986-1006
void adv_pos_half(Ora& o, double dt) {
    const double dth = 0.5 * dt;
    for (int64_t p = 0; p < o.nump; ++p) o.pxp[p] = o.px0[p] + dth * o.pu0[p];
}

// calcRhoTask, src/Hydro.cc:1034-1054
void calc_rho(Ora& o, const std::vector<double>& zv, std::vector<double>& zr) {
    for (int64_t z = 0; z < o.numz; ++z) zr[z] = o.zm[z] / zv[z];
}

// calcCrnrMassTask, src/Hydro.cc:1083-1117 (ascending s, one piece)
void calc_crnr_mass(Ora& o) {
    for (int64_t s = 0; s < o.nums; ++s) {
        const int64_t s3 = o.mapss3[s];
        const int64_t z = o.mapsz[s];
        const int64_t p = o.mapsp1[s];
        const double r = o.zrp[z];
        const double area = o.zareap[z];
        const double mf = o.smf[s];
        const double mf3 = o.smf[s3];
        const double mwt = r * area * 0.5 * (mf + mf3);
        o.pmaswt[p] += mwt;
    }
}

// sumCrnrForceTask, src/Hydro.cc:1159-1193
void sum_crnr_force(Ora& o) {
    for (int64_t s = 0; s < o.nums; ++s) {
        const int64_t s3 = o.mapss3[s];
        const int64_t p = o.mapsp1[s];
        const V2 cf = (o.sfp[s] + o.sfq[s] + o.sft[s]) - (o.sfp[s3] + o.sfq[s3] + o.sft[s3]);
        o.pf[p] = o.pf[p] + cf;
    }
}

// calcAccelTask, src/Hydro.cc:1235-1254
void calc_accel(Ora& o) {
    const double fuzz = 1.e-99;
    for (int64_t p = 0; p < o.nump; ++p) o.pap[p] = o.pf[p] / smax(o.pmaswt[p], fuzz);
}

// advPosFullTask, src/Hydro.cc:1281-1306
void adv_pos_full(Ora& o, double dt) {
    for (int64_t p = 0; p < o.nump; ++p) {
        const V2 x0 = o.px0[p], u0 = o.pu0[p], a = o.pap[p];
        const V2 u = u0 + dt * a;
        o.pu[p] = u;
        o.px[p] = x0 + dt * 0.5 * (u0 + u);
    }
}

// calcWorkTask, src/Hydro.cc:1339-1406
void calc_work(Ora& o, double dt) {
    for (int64_t z = 0; z < o.numz; ++z) o.zw[z] = 0.;
    const double dth = 0.5 * dt;
    for (int64_t s = 0; s < o.nums; ++s) {
        const int64_t p1 = o.mapsp1[s], p2 = o.mapsp2[s], z = o.mapsz[s];
        const V2 sftot = o.sfp[s] + o.sfq[s];
        const double sd1 = dot(sftot, (o.pu0[p1] + o.pu[p1]));
        const double sd2 = dot(-sftot, (o.pu0[p2] + o.pu[p2]));
        const double dwork = -dth * (sd1 * o.pxp[p1].x + sd2 * o.pxp[p2].x);
        o.zetot[z] += dwork;
        o.zw[z] += dwork;
    }
}

// calcWorkRateTask, src/Hydro.cc:1483-1510
void calc_work_rate(Ora& o, double dt) {
    double dtinv = 1. / dt;
    for (int64_t z = 0; z < o.numz; ++z) {
        const double dvol = o.zvol[z] - o.zvol0[z];
        o.zwrate[z] = (o.zw[z] + o.zp[z] * dvol) * dtinv;
    }
}

// calcEnergyTask, src/Hydro.cc:1545-1564
void calc_energy(Ora& o) {
    const double fuzz = 1.e-99;
    for (int64_t z = 0; z < o.numz; ++z) o.ze[z] = o.zetot[z] / (o.zm[z] + fuzz);
}

// calcDtNewTask, src/Hydro.cc:1591-1619
double calc_dt_new(Ora& o) {
    const double fuzz = 1.e-99;
    double dtnew = 1.e99;
    for (int64_t z = 0; z < o.numz; ++z) {
        const double cdu = smax(o.zdu[z], smax(o.zss[z], fuzz));
        const double zdthyd = o.zdl[z] * o.cfl / cdu;
        dtnew = (zdthyd < dtnew ? zdthyd : dtnew);
    }
    return dtnew;
}

// calcDvolTask, src/Hydro.cc:1622-1644
double calc_dvol(Ora& o) {
    double dvovmax = 1.e-99;
    for (int64_t z = 0; z < o.numz; ++z) {
        const double zdvov = std::fabs((o.zvol[z] - o.zvol0[z]) / o.zvol0[z]);
        dvovmax = (zdvov > dvovmax ? zdvov : dvovmax);
    }
    return dvovmax;
}

// calcDtTask, src/Hydro.cc:1647-1667
double calc_dt_hydro(double cflv, double dtlast, double dtnew, double dvovmax) {
    double dtrec = 1.e99;
    if (dtnew < dtrec) dtrec = dtnew;
    double dtnew2 = dtlast * cflv / dvovmax;
    if (dtnew2 < dtrec) dtrec = dtnew2;
    return dtrec;
}

// ---------------------------------------------------------------------------
// PolyGas (src/PolyGas.cc), TTS (src/TTS.cc).

// calcStateHalfTask, src/PolyGas.cc:63-108 (reads ZR, not ZRP)
void calc_state_half(Ora& o, double dt) {
    const double gamma = o.pg_gamma, ssmin = o.pg_ssmin;
    const double dth = 0.5 * dt;
    const double gm1 = gamma - 1.;
    const double ssmin2 = smax(ssmin * ssmin, 1.e-99);
    for (int64_t z = 0; z < o.numz; ++z) {
        const double r = o.zr[z];
        const double e = smax(o.ze[z], 0.);
        const double p = gm1 * r * e;
        const double pre = gm1 * e;
        const double per = gm1 * r;
        const double csqd = smax(ssmin2, pre + per * p / (r * r));
        const double ss = std::sqrt(csqd);
        const double minv = 1. / o.zm[z];
        const double volp = o.zvolp[z];
        const double vol0 = o.zvol0[z];
        const double wrate = o.zwrate[z];
        const double dv = (volp - vol0) * minv;
        const double bulk = r * csqd;
        const double denom = 1. + 0.5 * per * dv;
        const double src = wrate * dth * minv;
        o.zp[z] = p + (per * src - r * bulk * dv) / denom;
        o.zss[z] = ss;
    }
}

// calcForceTask, src/PolyGas.cc:164-185
void calc_force_pgas(Ora& o) {
    for (int64_t s = 0; s < o.nums; ++s) {
        const double p = o.zp[o.mapsz[s]];
        o.sfp[s] = -p * o.ssurfp[s];
    }
}

// TTS::calcForceTask, src/TTS.cc:55-104
void calc_force_tts(Ora& o) {
    const double alfa = o.tts_alfa, ssmin = o.tts_ssmin;
    for (int64_t s = 0; s < o.nums; ++s) {
        const int64_t z = o.mapsz[s];
        const double sarea = o.sareap[s];
        const double zarea = o.zareap[z];
        const double vfacinv = zarea / sarea;
        const double r = o.zrp[z];
        const double mf = o.smf[s];
        const double srho = r * mf * vfacinv;
        const double ss = o.zss[z];
        double sstmp = smax(ss, ssmin);
        sstmp = alfa * sstmp * sstmp;
        const double sdp = sstmp * (srho - r);
        o.sft[s] = -sdp * o.ssurfp[s];
    }
}

// ---------------------------------------------------------------------------
// QCS (src/QCS.cc).

// setCornerDivTask, src/QCS.cc:116-242
void set_corner_div(Ora& o) {
    for (int64_t z = 0; z < o.numz; ++z) o.zuc[z] = mk(0., 0.);
    for (int64_t s = 0; s < o.nums; ++s) {
        const int64_t z = o.mapsz[s];
        const int n = o.znump[z];
        o.zuc[z] = o.zuc[z] + o.pu0[o.mapsp1[s]] / (double)n;
    }
    for (int64_t c = 0; c < o.nums; ++c) {
        const int64_t s2 = c;
        const int64_t s = o.mapss3[s2];
        const int64_t z = o.mapsz[s];
        const int64_t p = o.mapsp2[s];
        const int64_t p1 = o.mapsp1[s];
        const int64_t p2 = o.mapsp2[s2];
        const V2 up0 = o.pu0[p];
        const V2 xp0 = o.pxp[p];
        const V2 up1 = 0.5 * (up0 + o.pu0[p2]);
        const V2 xp1 = o.exp_[s2];
        const V2 up2 = o.zuc[z];
        const V2 xp2 = o.zxp[z];
        const V2 up3 = 0.5 * (o.pu0[p1] + up0);
        const V2 xp3 = o.exp_[s];
        double cvolume = 0.5 * cross(xp2 - xp0, xp3 - xp1);
        o.carea[c] = cvolume;
        const V2 v1 = xp3 - xp0;
        const V2 v2 = xp1 - xp0;
        const double de1 = o.elen[s];
        const double de2 = o.elen[s2];
        const double minelen = smin(de1, de2);
        const double ccos = ((minelen < 1.e-12) ? 0. : 4. * dot(v1, v2) / (de1 * de2));
        o.ccos[c] = ccos;
        const double cdiv = (cross(up2 - up0, xp3 - xp1) - cross(up3 - up1, xp2 - xp0)) /
                            (2.0 * cvolume);
        o.cdiv[c] = cdiv;
        const V2 dxx1 = 0.5 * (xp1 + xp2 - xp0 - xp3);
        const V2 dxx2 = 0.5 * (xp2 + xp3 - xp0 - xp1);
        const double dx1 = length(dxx1);
        const double dx2 = length(dxx2);
        const V2 duav = 0.25 * (up0 + up1 + up2 + up3);
        const double test1 = std::fabs(dot(dxx1, duav) * dx2);
        const double test2 = std::fabs(dot(dxx2, duav) * dx1);
        const double num = (test1 > test2 ? dx1 : dx2);
        const double den = (test1 > test2 ? dx2 : dx1);
        const double r = num / den;
        double evol = std::sqrt(4.0 * cvolume * r);
        evol = smin(evol, 2.0 * minelen);
        const double dv1 = length2(up1 + up2 - up0 - up3);
        const double dv2 = length2(up2 + up3 - up0 - up1);
        double du = std::sqrt(smax(dv1, dv2));
        evol = (cdiv < 0.0 ? evol : 0.);
        du = (cdiv < 0.0 ? du : 0.);
        o.cevol[c] = evol;
        o.cdu[c] = du;
    }
}

// setQCnForceTask, src/QCS.cc:246-327
void set_qcn_force(Ora& o) {
    const double gammap1 = o.qgamma + 1.0;
    for (int64_t c = 0; c < o.nums; ++c) {
        const int64_t z = o.mapsz[c];
        const double cdu = o.cdu[c];
        const double ztmp2 = o.q2 * 0.25 * gammap1 * cdu;
        const double zss = o.zss[z];
        const double ztmp1 = o.q1 * zss;
        const double zkur = ztmp2 + std::sqrt(ztmp2 * ztmp2 + ztmp1 * ztmp1);
        const double zrp = o.zrp[z];
        const double cevol = o.cevol[c];
        const double crmu = zkur * zrp * cevol;
        const double cdiv = o.cdiv[c];
        o.crmu[c] = ((cdiv > 0.0) ? 0. : crmu);
    }
    for (int64_t c = 0; c < o.nums; ++c) {
        const int64_t s2 = c;
        const int64_t s = o.mapss3[s2];
        const int64_t p = o.mapsp2[s];
        const int64_t p1 = o.mapsp1[s];
        const int64_t p2 = o.mapsp2[s2];
        const double crmu = o.crmu[c];
        const V2 pu = o.pu0[p];
        const V2 pu1 = o.pu0[p1];
        o.cqe1[c] = crmu * (pu - pu1) / o.elen[s];
        const V2 pu2 = o.pu0[p2];
        o.cqe2[c] = crmu * (pu2 - pu) / o.elen[s2];
    }
}

// setForceTask, src/QCS.cc:331-380
void set_force_qcs(Ora& o) {
    for (int64_t c = 0; c < o.nums; ++c) {
        const double ccos = o.ccos[c];
        const double csin2 = 1.0 - ccos * ccos;
        o.cw[c] = ((csin2 < 1.e-4) ? 0. : o.carea[c] / csin2);
        o.ccos[c] = ((csin2 < 1.e-4) ? 0. : ccos);
    }
    for (int64_t s = 0; s < o.nums; ++s) {
        const int64_t c1 = s, c2 = o.mapss4[s];
        const double el = o.elen[s];
        const double cw1 = o.cw[c1], cw2 = o.cw[c2];
        const double ccos1 = o.ccos[c1], ccos2 = o.ccos[c2];
        const V2 cqe11 = o.cqe1[c1], cqe12 = o.cqe1[c2];
        const V2 cqe21 = o.cqe2[c1], cqe22 = o.cqe2[c2];
        o.sfq[s] = (cw1 * (cqe21 + ccos1 * cqe11) + cw2 * (cqe12 + ccos2 * cqe22)) / el;
    }
}

// setVelDiffTask, src/QCS.cc:384-448
void set_vel_diff(Ora& o) {
    for (int64_t z = 0; z < o.numz; ++z) o.ztmp[z] = 0.;
    for (int64_t s = 0; s < o.nums; ++s) {
        const int64_t p1 = o.mapsp1[s], p2 = o.mapsp2[s], z = o.mapsz[s];
        const V2 dx = o.pxp[p2] - o.pxp[p1];
        const V2 du = o.pu0[p2] - o.pu0[p1];
        const double lenx = o.elen[s];
        double dux = dot(du, dx);
        dux = (lenx > 0. ? std::fabs(dux) / lenx : 0.);
        o.ztmp[z] = smax(o.ztmp[z], dux);
    }
    for (int64_t z = 0; z < o.numz; ++z)
        o.zdu[z] = o.q1 * o.zss[z] + 2. * o.q2 * o.ztmp[z];
}

// ---------------------------------------------------------------------------
// HydroBC (src/HydroBC.cc).

// boundary point lists, src/HydroBC.cc:63-144,222-326 (one piece: private
// points in ascending order)
void make_bcs(Ora& o) {
    const double eps = 1.e-12;
    for (double b : o.bcx) {
        BC bc; bc.vfix = mk(1., 0.);
        for (int64_t p = 0; p < o.nump; ++p) if (std::fabs(o.px[p].x - b) < eps) bc.pts.push_back(p);
        o.bcs.push_back(bc);
    }
    for (double b : o.bcy) {
        BC bc; bc.vfix = mk(0., 1.);
        for (int64_t p = 0; p < o.nump; ++p) if (std::fabs(o.px[p].y - b) < eps) bc.pts.push_back(p);
        o.bcs.push_back(bc);
    }
}

// applyFixedBCTask, src/HydroBC.cc:150-183
void apply_fixed_bc(Ora& o) {
    for (const BC& bc : o.bcs) {
        for (int64_t p : bc.pts) {
            o.pu0[p] = project(o.pu0[p], bc.vfix);
            o.pf[p] = project(o.pf[p], bc.vfix);
        }
    }
}

// ---------------------------------------------------------------------------
// OpenMP variants used only for the multi-core CPU timing baseline; they
// mirror the reference's *OMPTask bodies (atomic accumulation), so their sums
// are order-nondeterministic exactly as the reference's are.

#ifdef _OPENMP
inline void atomic_add(double& lhs, double v) {
#pragma omp atomic
    lhs += v;
}
inline void atomic_min(double& lhs, double v) {
    // CAS loop as src/Mesh.hh:197-214
    int64_t* t = reinterpret_cast<int64_t*>(&lhs);
    union { int64_t i; double d; } ov, nv;
    do { ov.i = *t; nv.d = (ov.d < v) ? ov.d : v; }
    while (!__sync_bool_compare_and_swap(t, ov.i, nv.i));
}
inline void atomic_max(double& lhs, double v) {
    int64_t* t = reinterpret_cast<int64_t*>(&lhs);
    union { int64_t i; double d; } ov, nv;
    do { ov.i = *t; nv.d = (ov.d > v) ? ov.d : v; }
    while (!__sync_bool_compare_and_swap(t, ov.i, nv.i));
}
#endif

// ---------------------------------------------------------------------------
// doCycle, src/Hydro.cc:342-983, one step per reference launch group.

enum Step {
    ST_COPY_ZVOL0 = 0, ST_POINT_INIT, ST_CTRS_PRED, ST_VOLS_PRED, ST_SURFVECS,
    ST_EDGELEN, ST_CHARLEN, ST_RHO_PRED, ST_CRNRMASS, ST_STATEHALF, ST_FORCEPGAS,
    ST_FORCETTS, ST_CORNERDIV, ST_QCNFORCE, ST_FORCEQCS, ST_VELDIFF,
    ST_SUMCRNRFORCE, ST_FIXEDBC, ST_CHECK_PRED, ST_ACCEL_ADVFULL, ST_CTRS_CORR,
    ST_VOLS_CORR, ST_WORK, ST_WORKRATE, ST_ENERGY, ST_RHO_CORR, ST_DTNEW, ST_DVOL,
    ST_CALCDT, ST_CHECK_CORR, ST_NSTEPS
};

void omp_step(Ora& o, int st, double dt);

// returns 0, or 1 when a bad-side check fails
int do_step(Ora& o, int st, double dt) {
    if (o.use_omp) {
#ifdef _OPENMP
        switch (st) {
            case ST_CTRS_PRED: case ST_VOLS_PRED: case ST_SURFVECS: case ST_EDGELEN:
            case ST_CHARLEN: case ST_CRNRMASS: case ST_FORCEPGAS: case ST_FORCETTS:
            case ST_CORNERDIV: case ST_QCNFORCE: case ST_FORCEQCS: case ST_VELDIFF:
            case ST_SUMCRNRFORCE: case ST_CTRS_CORR: case ST_VOLS_CORR: case ST_WORK:
            case ST_STATEHALF: case ST_ACCEL_ADVFULL: case ST_POINT_INIT:
                omp_step(o, st, dt);
                return 0;
            default: break;
        }
#endif
    }
    switch (st) {
        case ST_COPY_ZVOL0: o.zvol0 = o.zvol; break;                    // Hydro.cc:361-368
        case ST_POINT_INIT:                                             // Hydro.cc:389-437
            o.px0 = o.px; o.pu0 = o.pu;
            for (int64_t p = 0; p < o.nump; ++p) { o.pmaswt[p] = 0.; o.pf[p] = mk(0., 0.); }
            adv_pos_half(o, dt);
            break;
        case ST_CTRS_PRED: calc_ctrs(o, o.pxp, o.exp_, o.zxp); break;   // Hydro.cc:439-464
        case ST_VOLS_PRED:                                              // Hydro.cc:466-493
            o.numsbad_pred = calc_vols(o, o.pxp, o.zxp, o.sareap, o.svolp, o.zareap, o.zvolp);
            break;
        case ST_SURFVECS: calc_surf_vecs(o); break;
        case ST_EDGELEN: calc_edge_len(o); break;
        case ST_CHARLEN: calc_char_len(o); break;
        case ST_RHO_PRED: calc_rho(o, o.zvolp, o.zrp); break;
        case ST_CRNRMASS: calc_crnr_mass(o); break;
        case ST_STATEHALF: calc_state_half(o, dt); break;
        case ST_FORCEPGAS: calc_force_pgas(o); break;
        case ST_FORCETTS: calc_force_tts(o); break;
        case ST_CORNERDIV: set_corner_div(o); break;
        case ST_QCNFORCE: set_qcn_force(o); break;
        case ST_FORCEQCS: set_force_qcs(o); break;
        case ST_VELDIFF: set_vel_diff(o); break;
        case ST_SUMCRNRFORCE: sum_crnr_force(o); break;
        case ST_FIXEDBC: apply_fixed_bc(o); break;
        case ST_CHECK_PRED: if (o.numsbad_pred > 0) return 1; break;   // Mesh.cc:1531-1550
        case ST_ACCEL_ADVFULL: calc_accel(o); adv_pos_full(o, dt); break;
        case ST_CTRS_CORR: calc_ctrs(o, o.px, o.ex, o.zx); break;       // Hydro.cc:860-870
        case ST_VOLS_CORR:                                              // Hydro.cc:872-883
            o.numsbad_corr = calc_vols(o, o.px, o.zx, o.sarea, o.svol, o.zarea, o.zvol);
            break;
        case ST_WORK: calc_work(o, dt); break;
        case ST_WORKRATE: calc_work_rate(o, dt); break;
        case ST_ENERGY: calc_energy(o); break;
        case ST_RHO_CORR: calc_rho(o, o.zvol, o.zr); break;
        case ST_DTNEW: o.dtnew = calc_dt_new(o); break;
        case ST_DVOL: o.dvovmax = calc_dvol(o); break;
        case ST_CALCDT: o.dtrec = calc_dt_hydro(o.cflv, dt, o.dtnew, o.dvovmax); break;
        case ST_CHECK_CORR: if (o.numsbad_corr > 0) return 1; break;
        default: break;
    }
    return 0;
}

#ifdef _OPENMP
// The reference's OMP task bodies (e.g. src/Mesh.cc:840-880, src/Hydro.cc
// *OMPTask) restated with their atomics.
void omp_step(Ora& o, int st, double dt) {
    const int64_t S = o.nums, Z = o.numz, P = o.nump;
    switch (st) {
    case ST_POINT_INIT: {
        const double dth = 0.5 * dt;
#pragma omp parallel for
        for (int64_t p = 0; p < P; ++p) {
            o.px0[p] = o.px[p]; o.pu0[p] = o.pu[p]; o.pmaswt[p] = 0.; o.pf[p] = mk(0., 0.);
            o.pxp[p] = o.px0[p] + dth * o.pu0[p];
        }
        break;
    }
    case ST_CTRS_PRED: case ST_CTRS_CORR: {
        const std::vector<V2>& px = (st == ST_CTRS_PRED) ? o.pxp : o.px;
        std::vector<V2>& ex = (st == ST_CTRS_PRED) ? o.exp_ : o.ex;
        std::vector<V2>& zx = (st == ST_CTRS_PRED) ? o.zxp : o.zx;
#pragma omp parallel for
        for (int64_t z = 0; z < Z; ++z) zx[z] = mk(0., 0.);
#pragma omp parallel for
        for (int64_t s = 0; s < S; ++s) {
            const V2 px1 = px[o.mapsp1[s]], px2 = px[o.mapsp2[s]];
            ex[s] = 0.5 * (px1 + px2);
            const int64_t z = o.mapsz[s];
            const V2 d = px1 / (double)o.znump[z];
            atomic_add(zx[z].x, d.x); atomic_add(zx[z].y, d.y);
        }
        break;
    }
    case ST_VOLS_PRED: case ST_VOLS_CORR: {
        const bool pr = (st == ST_VOLS_PRED);
        const std::vector<V2>& px = pr ? o.pxp : o.px;
        const std::vector<V2>& zx = pr ? o.zxp : o.zx;
        std::vector<double>& sa = pr ? o.sareap : o.sarea;
        std::vector<double>& sv = pr ? o.svolp : o.svol;
        std::vector<double>& za = pr ? o.zareap : o.zarea;
        std::vector<double>& zv = pr ? o.zvolp : o.zvol;
#pragma omp parallel for
        for (int64_t z = 0; z < Z; ++z) { za[z] = 0.; zv[z] = 0.; }
        const double third = 1. / 3.;
        int count = 0;
#pragma omp parallel for reduction(+:count)
        for (int64_t s = 0; s < S; ++s) {
            const V2 px1 = px[o.mapsp1[s]], px2 = px[o.mapsp2[s]];
            const int64_t z = o.mapsz[s];
            const V2 c = zx[z];
            const double a = 0.5 * cross(px2 - px1, c - px1);
            const double v = third * a * (px1.x + px2.x + c.x);
            sa[s] = a; sv[s] = v;
            atomic_add(za[z], a); atomic_add(zv[z], v);
            if (v <= 0.) count += 1;
        }
        if (pr) o.numsbad_pred = count; else o.numsbad_corr = count;
        break;
    }
    case ST_SURFVECS:
#pragma omp parallel for
        for (int64_t s = 0; s < S; ++s) o.ssurfp[s] = rotateCCW(o.exp_[s] - o.zxp[o.mapsz[s]]);
        break;
    case ST_EDGELEN:
#pragma omp parallel for
        for (int64_t s = 0; s < S; ++s) o.elen[s] = length(o.pxp[o.mapsp2[s]] - o.pxp[o.mapsp1[s]]);
        break;
    case ST_CHARLEN:
#pragma omp parallel for
        for (int64_t z = 0; z < Z; ++z) o.zdl[z] = 1.e99;
#pragma omp parallel for
        for (int64_t s = 0; s < S; ++s) {
            const int64_t z = o.mapsz[s];
            const double fac = (o.znump[z] == 3 ? 3. : 4.);
            atomic_min(o.zdl[z], fac * o.sareap[s] / o.elen[s]);
        }
        break;
    case ST_CRNRMASS:
#pragma omp parallel for
        for (int64_t s = 0; s < S; ++s) {
            const int64_t z = o.mapsz[s];
            const double mwt = o.zrp[z] * o.zareap[z] * 0.5 * (o.smf[s] + o.smf[o.mapss3[s]]);
            atomic_add(o.pmaswt[o.mapsp1[s]], mwt);
        }
        break;
    case ST_STATEHALF: {
        // same body as the serial variant, iterations are independent
        const double gamma = o.pg_gamma, ssmin = o.pg_ssmin, dth = 0.5 * dt, gm1 = gamma - 1.;
        const double ssmin2 = smax(ssmin * ssmin, 1.e-99);
#pragma omp parallel for
        for (int64_t z = 0; z < Z; ++z) {
            const double r = o.zr[z], e = smax(o.ze[z], 0.), p = gm1 * r * e;
            const double pre = gm1 * e, per = gm1 * r;
            const double csqd = smax(ssmin2, pre + per * p / (r * r));
            const double minv = 1. / o.zm[z];
            const double dv = (o.zvolp[z] - o.zvol0[z]) * minv;
            const double bulk = r * csqd, denom = 1. + 0.5 * per * dv;
            const double src = o.zwrate[z] * dth * minv;
            o.zp[z] = p + (per * src - r * bulk * dv) / denom;
            o.zss[z] = std::sqrt(csqd);
        }
        break;
    }
    case ST_FORCEPGAS:
#pragma omp parallel for
        for (int64_t s = 0; s < S; ++s) o.sfp[s] = -o.zp[o.mapsz[s]] * o.ssurfp[s];
        break;
    case ST_FORCETTS:
#pragma omp parallel for
        for (int64_t s = 0; s < S; ++s) {
            const int64_t z = o.mapsz[s];
            const double r = o.zrp[z];
            const double srho = r * o.smf[s] * (o.zareap[z] / o.sareap[s]);
            double t = smax(o.zss[z], o.tts_ssmin);
            t = o.tts_alfa * t * t;
            o.sft[s] = -(t * (srho - r)) * o.ssurfp[s];
        }
        break;
    case ST_CORNERDIV: case ST_QCNFORCE: case ST_FORCEQCS: {
        // corner loops are independent; the zuc sum uses atomics
        if (st == ST_CORNERDIV) {
#pragma omp parallel for
            for (int64_t z = 0; z < Z; ++z) o.zuc[z] = mk(0., 0.);
#pragma omp parallel for
            for (int64_t s = 0; s < S; ++s) {
                const int64_t z = o.mapsz[s];
                const V2 d = o.pu0[o.mapsp1[s]] / (double)o.znump[z];
                atomic_add(o.zuc[z].x, d.x); atomic_add(o.zuc[z].y, d.y);
            }
            // the per-corner part: reuse the serial body restricted to corners
            std::vector<V2> zuc_save = o.zuc;
            Ora* op = &o;
#pragma omp parallel for
            for (int64_t c = 0; c < S; ++c) {
                Ora& q = *op;
                const int64_t s2 = c, s = q.mapss3[s2], z = q.mapsz[s];
                const int64_t p = q.mapsp2[s], p1 = q.mapsp1[s], p2 = q.mapsp2[s2];
                const V2 up0 = q.pu0[p], xp0 = q.pxp[p];
                const V2 up1 = 0.5 * (up0 + q.pu0[p2]), xp1 = q.exp_[s2];
                const V2 up2 = zuc_save[z], xp2 = q.zxp[z];
                const V2 up3 = 0.5 * (q.pu0[p1] + up0), xp3 = q.exp_[s];
                const double cvolume = 0.5 * cross(xp2 - xp0, xp3 - xp1);
                q.carea[c] = cvolume;
                const V2 v1 = xp3 - xp0, v2 = xp1 - xp0;
                const double de1 = q.elen[s], de2 = q.elen[s2], minelen = smin(de1, de2);
                q.ccos[c] = ((minelen < 1.e-12) ? 0. : 4. * dot(v1, v2) / (de1 * de2));
                const double cdiv = (cross(up2 - up0, xp3 - xp1) - cross(up3 - up1, xp2 - xp0)) / (2.0 * cvolume);
                q.cdiv[c] = cdiv;
                const V2 dxx1 = 0.5 * (xp1 + xp2 - xp0 - xp3), dxx2 = 0.5 * (xp2 + xp3 - xp0 - xp1);
                const double dx1 = length(dxx1), dx2 = length(dxx2);
                const V2 duav = 0.25 * (up0 + up1 + up2 + up3);
                const double test1 = std::fabs(dot(dxx1, duav) * dx2), test2 = std::fabs(dot(dxx2, duav) * dx1);
                const double r = (test1 > test2 ? dx1 : dx2) / (test1 > test2 ? dx2 : dx1);
                double evol = smin(std::sqrt(4.0 * cvolume * r), 2.0 * minelen);
                double du = std::sqrt(smax(length2(up1 + up2 - up0 - up3), length2(up2 + up3 - up0 - up1)));
                q.cevol[c] = (cdiv < 0.0 ? evol : 0.);
                q.cdu[c] = (cdiv < 0.0 ? du : 0.);
            }
        } else if (st == ST_QCNFORCE) {
            const double gammap1 = o.qgamma + 1.0;
#pragma omp parallel for
            for (int64_t c = 0; c < S; ++c) {
                const int64_t z = o.mapsz[c];
                const double ztmp2 = o.q2 * 0.25 * gammap1 * o.cdu[c], ztmp1 = o.q1 * o.zss[z];
                const double zkur = ztmp2 + std::sqrt(ztmp2 * ztmp2 + ztmp1 * ztmp1);
                o.crmu[c] = ((o.cdiv[c] > 0.0) ? 0. : zkur * o.zrp[z] * o.cevol[c]);
            }
#pragma omp parallel for
            for (int64_t c = 0; c < S; ++c) {
                const int64_t s = o.mapss3[c], p = o.mapsp2[s];
                o.cqe1[c] = o.crmu[c] * (o.pu0[p] - o.pu0[o.mapsp1[s]]) / o.elen[s];
                o.cqe2[c] = o.crmu[c] * (o.pu0[o.mapsp2[c]] - o.pu0[p]) / o.elen[c];
            }
        } else {
#pragma omp parallel for
            for (int64_t c = 0; c < S; ++c) {
                const double ccos = o.ccos[c], csin2 = 1.0 - ccos * ccos;
                o.cw[c] = ((csin2 < 1.e-4) ? 0. : o.carea[c] / csin2);
                o.ccos[c] = ((csin2 < 1.e-4) ? 0. : ccos);
            }
#pragma omp parallel for
            for (int64_t s = 0; s < S; ++s) {
                const int64_t c2 = o.mapss4[s];
                o.sfq[s] = (o.cw[s] * (o.cqe2[s] + o.ccos[s] * o.cqe1[s]) +
                            o.cw[c2] * (o.cqe1[c2] + o.ccos[c2] * o.cqe2[c2])) / o.elen[s];
            }
        }
        break;
    }
    case ST_VELDIFF:
#pragma omp parallel for
        for (int64_t z = 0; z < Z; ++z) o.ztmp[z] = 0.;
#pragma omp parallel for
        for (int64_t s = 0; s < S; ++s) {
            const int64_t p1 = o.mapsp1[s], p2 = o.mapsp2[s];
            const double lenx = o.elen[s];
            double dux = dot(o.pu0[p2] - o.pu0[p1], o.pxp[p2] - o.pxp[p1]);
            dux = (lenx > 0. ? std::fabs(dux) / lenx : 0.);
            atomic_max(o.ztmp[o.mapsz[s]], dux);
        }
#pragma omp parallel for
        for (int64_t z = 0; z < Z; ++z) o.zdu[z] = o.q1 * o.zss[z] + 2. * o.q2 * o.ztmp[z];
        break;
    case ST_SUMCRNRFORCE:
#pragma omp parallel for
        for (int64_t s = 0; s < S; ++s) {
            const int64_t s3 = o.mapss3[s], p = o.mapsp1[s];
            const V2 cf = (o.sfp[s] + o.sfq[s] + o.sft[s]) - (o.sfp[s3] + o.sfq[s3] + o.sft[s3]);
            atomic_add(o.pf[p].x, cf.x); atomic_add(o.pf[p].y, cf.y);
        }
        break;
    case ST_ACCEL_ADVFULL:
#pragma omp parallel for
        for (int64_t p = 0; p < P; ++p) {
            o.pap[p] = o.pf[p] / smax(o.pmaswt[p], 1.e-99);
            const V2 u = o.pu0[p] + dt * o.pap[p];
            o.pu[p] = u;
            o.px[p] = o.px0[p] + dt * 0.5 * (o.pu0[p] + u);
        }
        break;
    case ST_WORK: {
        const double dth = 0.5 * dt;
#pragma omp parallel for
        for (int64_t z = 0; z < Z; ++z) o.zw[z] = 0.;
#pragma omp parallel for
        for (int64_t s = 0; s < S; ++s) {
            const int64_t p1 = o.mapsp1[s], p2 = o.mapsp2[s], z = o.mapsz[s];
            const V2 sftot = o.sfp[s] + o.sfq[s];
            const double sd1 = dot(sftot, (o.pu0[p1] + o.pu[p1]));
            const double sd2 = dot(-sftot, (o.pu0[p2] + o.pu[p2]));
            const double dwork = -dth * (sd1 * o.pxp[p1].x + sd2 * o.pxp[p2].x);
            atomic_add(o.zetot[z], dwork); atomic_add(o.zw[z], dwork);
        }
        break;
    }
    default: break;
    }
}
#endif

// calcGlobalDtTask, src/Driver.cc:213-245
double calc_global_dt(const Ora& o) {
    double dt = o.dtmax;
    if (o.cycle == 0) {
        if (o.dtinit < dt) dt = o.dtinit;
    } else {
        const double dtrecover = o.dtfac * o.dt;
        if (dtrecover < dt) dt = dtrecover;
    }
    if ((o.tstop - o.time) < dt) dt = o.tstop - o.time;
    if (o.cycle > 0) {
        if (o.dtrec < dt) dt = o.dtrec;
    }
    return dt;
}

bool setup(Ora& o, const char* text) {
    Deck d;
    if (!d.parse(text)) { o.err = d.err; return false; }
    o.cstop = d.geti("cstop", 999999);
    o.tstop = d.getd("tstop", 1.e99);
    if (o.cstop == 999999 && o.tstop == 1.e99) { o.err = "Must specify either cstop or tstop"; return false; }
    o.dtmax = d.getd("dtmax", 1.e99);
    o.dtinit = d.getd("dtinit", 1.e99);
    o.dtfac = d.getd("dtfac", 1.2);
    o.dtreport = d.geti("dtreport", 10);
    o.subregion = d.getl("subregion");
    if (!o.subregion.empty() && o.subregion.size() != 4) { o.err = "subregion must have 4 entries"; return false; }
    // GenMesh ctor, src/GenMesh.cc:87-129
    o.meshtype = d.gets("meshtype", "");
    if (o.meshtype.empty()) { o.err = "must specify meshtype"; return false; }
    if (o.meshtype != "pie" && o.meshtype != "rect" && o.meshtype != "hex") { o.err = "invalid meshtype " + o.meshtype; return false; }
    std::vector<double> params = d.getl("meshparams");
    if (params.empty()) { o.err = "must specify meshparams"; return false; }
    if (params.size() > 4) { o.err = "meshparams must have <= 4 values"; return false; }
    o.nzx = (int64_t)params[0];
    o.nzy = (params.size() >= 2 ? (int64_t)params[1] : o.nzx);
    if (o.meshtype != "pie") o.lenx = (params.size() >= 3 ? params[2] : 1.0);
    else o.lenx = (params.size() >= 3 ? params[2] : 90.0) * M_PI / 180.0;
    o.leny = (params.size() >= 4 ? params[3] : 1.0);
    if (o.nzx <= 0 || o.nzy <= 0 || o.lenx <= 0. || o.leny <= 0.) { o.err = "meshparams values must be positive"; return false; }
    if (o.meshtype == "pie" && o.lenx >= 2. * M_PI) { o.err = "meshparams theta must be < 360"; return false; }
    // Hydro ctor, src/Hydro.cc:200-228, and package ctors
    o.cfl = d.getd("cfl", 0.6); o.cflv = d.getd("cflv", 0.1);
    o.rinit = d.getd("rinit", 1.); o.einit = d.getd("einit", 0.);
    o.rinitsub = d.getd("rinitsub", 1.); o.einitsub = d.getd("einitsub", 0.);
    o.uinitradial = d.getd("uinitradial", 0.);
    o.bcx = d.getl("bcx"); o.bcy = d.getl("bcy");
    o.pg_gamma = d.getd("gamma", 5. / 3.); o.pg_ssmin = d.getd("ssmin", 0.);
    o.tts_alfa = d.getd("alfa", 0.5); o.tts_ssmin = d.getd("ssmin", 0.);
    o.qgamma = d.getd("qgamma", 5. / 3.); o.q1 = d.getd("q1", 0.); o.q2 = d.getd("q2", 2.);

    if (!gen_mesh(o)) return false;
    const int64_t P = o.nump, Z = o.numz, S = o.nums;
    for (auto* v : {&o.pxp, &o.px0, &o.pu, &o.pu0, &o.pf, &o.pap}) v->assign(P, mk(0., 0.));
    o.pmaswt.assign(P, 0.);
    for (auto* v : {&o.zx, &o.zxp, &o.zuc}) v->assign(Z, mk(0., 0.));
    for (auto* v : {&o.zarea, &o.zvol, &o.zareap, &o.zvolp, &o.zvol0, &o.zdl, &o.zm, &o.zr, &o.zrp,
                    &o.ze, &o.zetot, &o.zw, &o.zwrate, &o.zp, &o.zss, &o.zdu, &o.ztmp})
        v->assign(Z, 0.);
    for (auto* v : {&o.ex, &o.exp_, &o.ssurfp, &o.sfp, &o.sfq, &o.sft, &o.cqe1, &o.cqe2}) v->assign(S, mk(0., 0.));
    for (auto* v : {&o.sarea, &o.svol, &o.sareap, &o.svolp, &o.elen, &o.smf, &o.carea, &o.cevol,
                    &o.cdu, &o.cdiv, &o.ccos, &o.crmu, &o.cw})
        v->assign(S, 0.);

    // Mesh::init geometry, src/Mesh.cc:616-623
    calc_ctrs(o, o.px, o.ex, o.zx);
    int bad = calc_vols(o, o.px, o.zx, o.sarea, o.svol, o.zarea, o.zvol);
    if (bad > 0) {
        char b[128]; snprintf(b, sizeof b, "Error: %d negative side volumes during initialization", bad);
        o.err = b; return false;
    }
    calc_side_fracs(o);
    make_bcs(o);

    // Hydro::init, src/Hydro.cc:241-339
    for (int64_t z = 0; z < Z; ++z) { o.zr[z] = o.rinit; o.ze[z] = o.einit; o.zwrate[z] = 0.; }
    if (!o.subregion.empty()) {
        // initSubrgnTask, src/Hydro.cc:1738-1760
        const double eps = 1.e-12;
        const std::vector<double>& sr = o.subregion;
        for (int64_t z = 0; z < Z; ++z) {
            const V2 c = o.zx[z];
            if (c.x > (sr[0] - eps) && c.x < (sr[1] + eps) && c.y > (sr[2] - eps) && c.y < (sr[3] + eps)) {
                o.zr[z] = o.rinitsub; o.ze[z] = o.einitsub;
            }
        }
    }
    // initHydroTask, src/Hydro.cc:1763-1784
    for (int64_t z = 0; z < Z; ++z) {
        const double zm = o.zr[z] * o.zvol[z];
        o.zm[z] = zm;
        o.zetot[z] = o.ze[z] * zm;
    }
    if (o.uinitradial != 0.) {
        // initRadialVelTask, src/Hydro.cc:1787-1807
        const double eps = 1.e-12;
        for (int64_t p = 0; p < P; ++p) {
            const V2 x = o.px[p];
            const double pmag = length(x);
            if (pmag > eps) o.pu[p] = o.uinitradial * x / pmag;
            else o.pu[p] = mk(0., 0.);
        }
    } else {
        for (int64_t p = 0; p < P; ++p) o.pu[p] = mk(0., 0.);
    }
    return true;
}

// one cycle of Driver::run (src/Driver.cc:127-158) with PENNANT termination
// semantics; returns 0 on success, 1 on a bad-side error.
int run_cycle(Ora& o) {
    o.dt = calc_global_dt(o);
    for (int st = 0; st < ST_NSTEPS; ++st) {
        if (do_step(o, st, o.dt)) {
            char b[128];
            snprintf(b, sizeof b, "Error: %d negative side volumes on cycle %d",
                     st == ST_CHECK_PRED ? o.numsbad_pred : o.numsbad_corr, o.cycle);
            o.err = b;
            return 1;
        }
    }
    o.time = o.time + o.dt;   // updateTimeTask, src/Driver.cc:266-275
    o.cycle = o.cycle + 1;    // updateCycleTask, src/Driver.cc:288-295
    return 0;
}

struct FieldRef { const char* name; int kind; void* vec; };  // kind 1: double, 2: V2

std::vector<FieldRef> field_table(Ora& o) {
    return {
        {"px", 2, &o.px}, {"pxp", 2, &o.pxp}, {"px0", 2, &o.px0}, {"pu", 2, &o.pu},
        {"pu0", 2, &o.pu0}, {"pf", 2, &o.pf}, {"pap", 2, &o.pap}, {"pmaswt", 1, &o.pmaswt},
        {"zx", 2, &o.zx}, {"zxp", 2, &o.zxp}, {"zuc", 2, &o.zuc},
        {"zarea", 1, &o.zarea}, {"zvol", 1, &o.zvol}, {"zareap", 1, &o.zareap},
        {"zvolp", 1, &o.zvolp}, {"zvol0", 1, &o.zvol0}, {"zdl", 1, &o.zdl}, {"zm", 1, &o.zm},
        {"zr", 1, &o.zr}, {"zrp", 1, &o.zrp}, {"ze", 1, &o.ze}, {"zetot", 1, &o.zetot},
        {"zw", 1, &o.zw}, {"zwrate", 1, &o.zwrate}, {"zp", 1, &o.zp}, {"zss", 1, &o.zss},
        {"zdu", 1, &o.zdu}, {"ztmp", 1, &o.ztmp},
        {"ex", 2, &o.ex}, {"exp", 2, &o.exp_}, {"ssurfp", 2, &o.ssurfp}, {"sfp", 2, &o.sfp},
        {"sfq", 2, &o.sfq}, {"sft", 2, &o.sft}, {"cqe1", 2, &o.cqe1}, {"cqe2", 2, &o.cqe2},
        {"sarea", 1, &o.sarea}, {"svol", 1, &o.svol}, {"sareap", 1, &o.sareap},
        {"svolp", 1, &o.svolp}, {"elen", 1, &o.elen}, {"smf", 1, &o.smf}, {"carea", 1, &o.carea},
        {"cevol", 1, &o.cevol}, {"cdu", 1, &o.cdu}, {"cdiv", 1, &o.cdiv}, {"ccos", 1, &o.ccos},
        {"crmu", 1, &o.crmu}, {"cw", 1, &o.cw},
    };
}

}  // namespace

// ---------------------------------------------------------------------------
// C API used by the tests (ctypes).
extern "C" {

void* ora_create(const char* deck_text, char* errbuf, int errlen) {
    Ora* o = new Ora();
    if (!setup(*o, deck_text)) {
        if (errbuf && errlen > 0) snprintf(errbuf, errlen, "%s", o->err.c_str());
        delete o;
        return nullptr;
    }
    return o;
}

void ora_destroy(void* h) { delete static_cast<Ora*>(h); }

const char* ora_error(void* h) { return static_cast<Ora*>(h)->err.c_str(); }

void ora_set_omp(void* h, int on) {
#ifdef _OPENMP
    static_cast<Ora*>(h)->use_omp = (on != 0);
#else
    (void)h; (void)on;
#endif
}

int ora_omp_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// Runs up to maxcycles cycles, stopping at cstop or tstop.  Returns the
// number of cycles run, or -1 on a bad-side error (message in ora_error).
int ora_run(void* h, int maxcycles) {
    Ora& o = *static_cast<Ora*>(h);
    int n = 0;
    while (n < maxcycles && o.cycle < o.cstop && o.time < o.tstop) {
        if (run_cycle(o)) return -1;
        ++n;
    }
    return n;
}

int ora_done(void* h) {
    Ora& o = *static_cast<Ora*>(h);
    return (o.cycle >= o.cstop || o.time >= o.tstop) ? 1 : 0;
}

// Per-step access for task-level parity tests: ora_begin_cycle computes the
// cycle's dt (calcGlobalDt); ora_step runs one doCycle step with that dt;
// ora_end_cycle advances time and cycle.
double ora_begin_cycle(void* h) {
    Ora& o = *static_cast<Ora*>(h);
    o.dt = calc_global_dt(o);
    return o.dt;
}
int ora_step(void* h, int step) {
    Ora& o = *static_cast<Ora*>(h);
    if (step < 0 || step >= ST_NSTEPS) return -2;
    return do_step(o, step, o.dt);
}
void ora_end_cycle(void* h) {
    Ora& o = *static_cast<Ora*>(h);
    o.time = o.time + o.dt;
    o.cycle = o.cycle + 1;
}
int ora_nsteps(void) { return ST_NSTEPS; }

void ora_sizes(void* h, int64_t* out3) {
    Ora& o = *static_cast<Ora*>(h);
    out3[0] = o.nump; out3[1] = o.numz; out3[2] = o.nums;
}

// scalars: time, dt, dtrec, cycle, dtnew, dvovmax, numsbad_pred, numsbad_corr
void ora_scalars(void* h, double* out8) {
    Ora& o = *static_cast<Ora*>(h);
    out8[0] = o.time; out8[1] = o.dt; out8[2] = o.dtrec; out8[3] = o.cycle;
    out8[4] = o.dtnew; out8[5] = o.dvovmax; out8[6] = o.numsbad_pred; out8[7] = o.numsbad_corr;
}
void ora_set_scalars(void* h, const double* in8) {
    Ora& o = *static_cast<Ora*>(h);
    o.time = in8[0]; o.dt = in8[1]; o.dtrec = in8[2]; o.cycle = (int)in8[3];
    o.dtnew = in8[4]; o.dvovmax = in8[5];
}

// Field storage: returns a pointer to `count` values with `ncomp` doubles
// each (interleaved x,y for vector fields), or null for an unknown name.
double* ora_field(void* h, const char* name, int64_t* count, int* ncomp) {
    Ora& o = *static_cast<Ora*>(h);
    for (const FieldRef& f : field_table(o)) {
        if (std::strcmp(f.name, name) == 0) {
            if (f.kind == 1) {
                auto* v = static_cast<std::vector<double>*>(f.vec);
                *count = (int64_t)v->size(); *ncomp = 1; return v->data();
            }
            auto* v = static_cast<std::vector<V2>*>(f.vec);
            *count = (int64_t)v->size(); *ncomp = 2; return reinterpret_cast<double*>(v->data());
        }
    }
    return nullptr;
}

// maps: "mapsp1", "mapsp2", "mapsz", "mapss3", "mapss4" (int64), "znump" (int32)
const void* ora_map(void* h, const char* name, int64_t* count, int* width) {
    Ora& o = *static_cast<Ora*>(h);
    struct { const char* n; std::vector<int64_t>* v; } m[] = {
        {"mapsp1", &o.mapsp1}, {"mapsp2", &o.mapsp2}, {"mapsz", &o.mapsz},
        {"mapss3", &o.mapss3}, {"mapss4", &o.mapss4}};
    for (auto& e : m) if (std::strcmp(e.n, name) == 0) { *count = (int64_t)e.v->size(); *width = 8; return e.v->data(); }
    if (std::strcmp(name, "znump") == 0) { *count = (int64_t)o.znump.size(); *width = 4; return o.znump.data(); }
    return nullptr;
}

int ora_num_bcs(void* h) { return (int)static_cast<Ora*>(h)->bcs.size(); }
const int64_t* ora_bc(void* h, int i, int64_t* count, double* vfix2) {
    Ora& o = *static_cast<Ora*>(h);
    const BC& b = o.bcs[i];
    *count = (int64_t)b.pts.size(); vfix2[0] = b.vfix.x; vfix2[1] = b.vfix.y;
    return b.pts.data();
}

// Stand-alone scalar restatements for brute-force dt tests.
double ora_calc_dt_hydro(double cflv, double dtlast, double dtnew, double dvovmax) {
    return calc_dt_hydro(cflv, dtlast, dtnew, dvovmax);
}
double ora_calc_global_dt(double dtmax, double dtinit, double dtfac, double tstop,
                          int cycle, double time, double dtlast, double dtrec) {
    Ora o;
    o.dtmax = dtmax; o.dtinit = dtinit; o.dtfac = dtfac; o.tstop = tstop;
    o.cycle = cycle; o.time = time; o.dt = dtlast; o.dtrec = dtrec;
    return calc_global_dt(o);
}

// <probname>.xy writer, src/WriteXY.cc:28-53
int ora_write_xy(void* h, const char* path) {
    Ora& o = *static_cast<Ora*>(h);
    FILE* f = std::fopen(path, "w");
    if (!f) return -1;
    const std::vector<double>* arrs[3] = {&o.zr, &o.ze, &o.zp};
    const char* hdr[3] = {"#  zr", "#  ze", "#  zp"};
    for (int k = 0; k < 3; ++k) {
        std::fprintf(f, "%s\n", hdr[k]);
        for (int64_t z = 0; z < o.numz; ++z) std::fprintf(f, "%5lld%18.8e\n", (long long)(z + 1), (*arrs[k])[z]);
    }
    std::fclose(f);
    return 0;
}

}  // extern "C"
