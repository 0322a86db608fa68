// Fused PENNANT cycle for gfx950.
//
// The reference issues ~36 launches per cycle (src/Hydro.cc:342-983), each a
// streaming pass over sides, zones or points with intermediate arrays in
// memory.  Every task from advPosHalf to setVelDiff only couples a zone to
// its own sides and corners, so the whole predictor half of doCycle is one
// zone-local computation; the only cross-zone couplings are the two point
// sums (corner mass, corner force) and the corrector's point positions.  The
// fused cycle is therefore three passes plus a one-block finalize:
//
//   K1 zone pass   advPosHalf (per zone corner), calcCtrs, calcVols,
//                  calcSurfVecs, calcEdgeLen, calcCharLen, calcRho,
//                  calcCrnrMass (per-corner mass), calcStateHalf,
//                  calcForcePGas, calcForceTTS, setCornerDiv, setQCnForce,
//                  setForceQCS, setVelDiff, the corner force of
//                  sumCrnrForce, calcDtNew's per-zone dt
//   K2 point pass  corner-mass and corner-force gathers (ascending side
//                  order), applyFixedBC, calcAccel, advPosFull
//   K3 zone pass   calcCtrs, calcVols (corrector), calcWork, calcWorkRate,
//                  calcEnergy, calcRho, calcDvol's per-zone ratio
//   K4 finalize    dt min, dvol max and bad-side sums over block partials,
//                  calcDt, checkBadSides, time/cycle update, calcGlobalDt
//
// Every expression is evaluated exactly as in the task kernels (and the
// reference's serial bodies), so the end-of-cycle state is bit-identical to
// the task-by-task path.  Each kernel starts by reading the device `done`
// flag, which plays the role of the reference's cycle predicate
// (src/Driver.cc:140-144): launches after the stop condition do nothing, so
// the host can queue many cycles without reading anything back.
//
// Zones are processed one per thread, templated on the zone's side count NS
// so that all per-corner values live in registers; the host launches one
// kernel per run of consecutive zones with equal side count (rect: one run of
// quads; pie: a run of triangles, then quads).
#include "common.h"
#include "problem.h"

using namespace pn;

namespace {

struct FusedArgs {
    // mesh
    const int32_t* mapsp1;
    const int32_t* pcbeg;
    const int32_t* pcside;
    const double* smf;
    int nump;
    // state (in place)
    double* pxx; double* pxy; double* pux; double* puy;
    double* zr; double* ze; double* zetot; double* zwrate; double* zvol; double* zp;
    const double* zm;
    // per-side intermediates
    double* cfx; double* cfy; double* cmass; double* sfwx; double* sfwy;
    // per-point intermediates
    double* vsx; double* vsy; double* pxpx;
    const unsigned char* pbc;
    // partials
    double* part_dt; double* part_dv; int* part_bp; int* part_bc;
    Ctl* ctl;
    // physics parameters
    double gamma, ssmin, alfa, qgamma, q1, q2, cfl;
    int nbcs;
    double bcvx[8], bcvy[8];
};

template <int NS>
__device__ __forceinline__ int nxt(int k) { return (k + 1 == NS) ? 0 : k + 1; }
template <int NS>
__device__ __forceinline__ int prv(int k) { return (k == 0) ? NS - 1 : k - 1; }

// K1: predictor geometry, EOS, forces, QCS for one zone per thread.
template <int NS>
__global__ void __launch_bounds__(PN_BLOCK) k_fused_zone_pred(FusedArgs a, int zbeg, int zend,
                                                              int sbeg, int boff) {
    if (a.ctl->done) return;
    const int z = zbeg + blockIdx.x * PN_BLOCK + threadIdx.x;
    const bool live = z < zend;
    double zdthyd = 1.e99;
    int bad = 0;
    if (live) {
        const double dt = a.ctl->dt;
        const double dth = 0.5 * dt;
        const int sb = sbeg + (z - zbeg) * NS;
        const double n = (double)NS;
        d2 u0[NS], xp[NS];
        double mf[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int p = a.mapsp1[sb + k];
            const d2 x0 = mk(a.pxx[p], a.pxy[p]);
            u0[k] = mk(a.pux[p], a.puy[p]);
            xp[k] = add(x0, scl(dth, u0[k]));                 // advPosHalf
            mf[k] = a.smf[sb + k];
        }
        // calcCtrs
        d2 ex[NS];
        d2 zx = mk(0., 0.);
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            ex[k] = scl(0.5, add(xp[k], xp[nxt<NS>(k)]));
            zx = add(zx, dvs(xp[k], n));
        }
        // calcVols, calcEdgeLen, calcCharLen
        const double third = 1. / 3.;
        double sa[NS], el[NS];
        double zarea = 0., zvolp = 0.;
        const double fac = (NS == 3 ? 3. : 4.);
        double zdl = 1.e99;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const d2 p1 = xp[k], p2 = xp[nxt<NS>(k)];
            sa[k] = 0.5 * cross(sub(p2, p1), sub(zx, p1));
            const double sv = third * sa[k] * (p1.x + p2.x + zx.x);
            zarea += sa[k];
            zvolp += sv;
            if (sv <= 0.) bad += 1;
            el[k] = len(sub(p2, p1));
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) zdl = smin(zdl, fac * sa[k] / el[k]);
        // calcRho (pred)
        const double zm = a.zm[z];
        const double zrp = zm / zvolp;
        // calcStateHalf (reads ZR, ZE, ZWRATE of the previous cycle, ZVOL0 = ZVOL)
        const double gm1 = a.gamma - 1.;
        const double ssmin2 = smax(a.ssmin * a.ssmin, 1.e-99);
        const double r = a.zr[z];
        const double e = smax(a.ze[z], 0.);
        const double p = gm1 * r * e;
        const double pre = gm1 * e;
        const double per = gm1 * r;
        const double csqd = smax(ssmin2, pre + per * p / (r * r));
        const double zss = __builtin_sqrt(csqd);
        const double minv = 1. / zm;
        const double dv = (zvolp - a.zvol[z]) * minv;
        const double bulk = r * csqd;
        const double denom = 1. + 0.5 * per * dv;
        const double src = a.zwrate[z] * dth * minv;
        const double zp = p + (per * src - r * bulk * dv) / denom;
        a.zp[z] = zp;
        // calcSurfVecs, calcForcePGas, calcForceTTS
        d2 tot[NS], sfw[NS];
        double sstmp = smax(zss, a.ssmin);
        sstmp = a.alfa * sstmp * sstmp;
        d2 sfp[NS], sft[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const d2 ss = rot_ccw(sub(ex[k], zx));
            sfp[k] = scl(-zp, ss);
            const double vfacinv = zarea / sa[k];
            const double srho = zrp * mf[k] * vfacinv;
            const double sdp = sstmp * (srho - zrp);
            sft[k] = scl(-sdp, ss);
        }
        // setCornerDiv
        d2 zuc = mk(0., 0.);
#pragma unroll
        for (int k = 0; k < NS; ++k) zuc = add(zuc, dvs(u0[k], n));
        double carea[NS], ccos[NS], cdiv[NS], cevol[NS], cdu[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int km = prv<NS>(k), kp = nxt<NS>(k);
            const d2 up0 = u0[k];
            const d2 xp0 = xp[k];
            const d2 up1 = scl(0.5, add(up0, u0[kp]));
            const d2 xp1 = ex[k];
            const d2 up2 = zuc;
            const d2 xp2 = zx;
            const d2 up3 = scl(0.5, add(u0[km], up0));
            const d2 xp3 = ex[km];
            const double cvolume = 0.5 * cross(sub(xp2, xp0), sub(xp3, xp1));
            carea[k] = cvolume;
            const d2 v1 = sub(xp3, xp0);
            const d2 v2 = sub(xp1, xp0);
            const double de1 = el[km];
            const double de2 = el[k];
            const double minelen = smin(de1, de2);
            ccos[k] = ((minelen < 1.e-12) ? 0. : 4. * dot(v1, v2) / (de1 * de2));
            const double div = (cross(sub(up2, up0), sub(xp3, xp1)) - cross(sub(up3, up1), sub(xp2, xp0))) /
                               (2.0 * cvolume);
            cdiv[k] = div;
            const d2 dxx1 = scl(0.5, sub(sub(add(xp1, xp2), xp0), xp3));
            const d2 dxx2 = scl(0.5, sub(sub(add(xp2, xp3), xp0), xp1));
            const double dx1 = len(dxx1);
            const double dx2 = len(dxx2);
            const d2 duav = scl(0.25, add(add(add(up0, up1), up2), up3));
            const double test1 = fabs(dot(dxx1, duav) * dx2);
            const double test2 = fabs(dot(dxx2, duav) * dx1);
            const double num = (test1 > test2 ? dx1 : dx2);
            const double den = (test1 > test2 ? dx2 : dx1);
            const double rr = num / den;
            double evol = __builtin_sqrt(4.0 * cvolume * rr);
            evol = smin(evol, 2.0 * minelen);
            const double dv1 = len2(sub(sub(add(up1, up2), up0), up3));
            const double dv2 = len2(sub(sub(add(up2, up3), up0), up1));
            double du = __builtin_sqrt(smax(dv1, dv2));
            cevol[k] = (div < 0.0 ? evol : 0.);
            cdu[k] = (div < 0.0 ? du : 0.);
        }
        // setQCnForce
        const double gammap1 = a.qgamma + 1.0;
        const double ztmp1 = a.q1 * zss;
        d2 cqe1[NS], cqe2[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int km = prv<NS>(k), kp = nxt<NS>(k);
            const double ztmp2 = a.q2 * 0.25 * gammap1 * cdu[k];
            const double zkur = ztmp2 + __builtin_sqrt(ztmp2 * ztmp2 + ztmp1 * ztmp1);
            const double mu = zkur * zrp * cevol[k];
            const double cm = ((cdiv[k] > 0.0) ? 0. : mu);
            cqe1[k] = dvs(scl(cm, sub(u0[k], u0[km])), el[km]);
            cqe2[k] = dvs(scl(cm, sub(u0[kp], u0[k])), el[k]);
        }
        // setForceQCS
        double cw[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const double cc = ccos[k];
            const double csin2 = 1.0 - cc * cc;
            cw[k] = ((csin2 < 1.e-4) ? 0. : carea[k] / csin2);
            ccos[k] = ((csin2 < 1.e-4) ? 0. : cc);
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int kp = nxt<NS>(k);
            const d2 f = add(scl(cw[k], add(cqe2[k], scl(ccos[k], cqe1[k]))),
                             scl(cw[kp], add(cqe1[kp], scl(ccos[kp], cqe2[kp]))));
            const d2 sfq = dvs(f, el[k]);
            sfw[k] = add(sfp[k], sfq);                       // calcWork's sf + sf2
            tot[k] = add(add(sfp[k], sfq), sft[k]);          // sumCrnrForce's sfp+sfq+sft
        }
        // setVelDiff
        double ztmp = 0.;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int kp = nxt<NS>(k);
            const d2 dx = sub(xp[kp], xp[k]);
            const d2 du = sub(u0[kp], u0[k]);
            const double lenx = el[k];
            double dux = dot(du, dx);
            dux = (lenx > 0. ? fabs(dux) / lenx : 0.);
            ztmp = smax(ztmp, dux);
        }
        const double zdu = a.q1 * zss + 2. * a.q2 * ztmp;
        // calcDtNew, per zone
        const double fuzz = 1.e-99;
        const double cdt = smax(zdu, smax(zss, fuzz));
        zdthyd = zdl * a.cfl / cdt;
        // per-corner outputs: corner force and corner mass at corner k (point
        // mapsp1[sb+k]), and the work force of side k
        const double mzz = zrp * zarea;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int km = prv<NS>(k);
            const d2 cf = sub(tot[k], tot[km]);
            a.cfx[sb + k] = cf.x;
            a.cfy[sb + k] = cf.y;
            a.cmass[sb + k] = mzz * 0.5 * (mf[k] + mf[km]);
            a.sfwx[sb + k] = sfw[k].x;
            a.sfwy[sb + k] = sfw[k].y;
        }
        // NaN never replaces the running minimum (src/Hydro.cc:1615)
        if (!(zdthyd < 1.e99)) zdthyd = 1.e99;
    }
    const double m = block_min(zdthyd);
    const int b = block_sum(bad);
    if (threadIdx.x == 0) {
        a.part_dt[boff + blockIdx.x] = m;
        a.part_bp[boff + blockIdx.x] = b;
    }
}

// K2: point gathers, fixed BCs, acceleration and position update.
__global__ void __launch_bounds__(PN_BLOCK) k_fused_point(FusedArgs a) {
    if (a.ctl->done) return;
    const int p = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (p >= a.nump) return;
    const double dt = a.ctl->dt;
    const double dth = 0.5 * dt;
    double mass = 0.;
    d2 f = mk(0., 0.);
    const int kb = a.pcbeg[p], ke = a.pcbeg[p + 1];
    for (int k = kb; k < ke; ++k) {
        const int s = a.pcside[k];
        mass += a.cmass[s];
        f = add(f, mk(a.cfx[s], a.cfy[s]));
    }
    const d2 x0 = mk(a.pxx[p], a.pxy[p]);
    const d2 uin = mk(a.pux[p], a.puy[p]);
    d2 u0 = uin;
    const unsigned bits = a.pbc[p];
    if (bits) {
        for (int b = 0; b < a.nbcs; ++b) {
            if (bits & (1u << b)) {
                const d2 v = mk(a.bcvx[b], a.bcvy[b]);
                u0 = project(u0, v);
                f = project(f, v);
            }
        }
    }
    const double fuzz = 1.e-99;
    const d2 acc = dvs(f, smax(mass, fuzz));
    const d2 u = add(u0, scl(dt, acc));
    const d2 x = add(x0, scl(dt * 0.5, add(u0, u)));
    a.pxpx[p] = x0.x + dth * uin.x;          // pxp.x for calcWork
    const d2 vs = add(u0, u);                // pu0 + pu for calcWork
    a.vsx[p] = vs.x;
    a.vsy[p] = vs.y;
    a.pxx[p] = x.x;
    a.pxy[p] = x.y;
    a.pux[p] = u.x;
    a.puy[p] = u.y;
}

// K3: corrector geometry, work, work rate, energy, density, dvol.
template <int NS>
__global__ void __launch_bounds__(PN_BLOCK) k_fused_zone_corr(FusedArgs a, int zbeg, int zend,
                                                              int sbeg, int boff) {
    if (a.ctl->done) return;
    const int z = zbeg + blockIdx.x * PN_BLOCK + threadIdx.x;
    double dvov = 1.e-99;
    int bad = 0;
    if (z < zend) {
        const double dt = a.ctl->dt;
        const double dth = 0.5 * dt;
        const int sb = sbeg + (z - zbeg) * NS;
        const double n = (double)NS;
        int pp[NS];
        d2 x[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            pp[k] = a.mapsp1[sb + k];
            x[k] = mk(a.pxx[pp[k]], a.pxy[pp[k]]);
        }
        d2 zx = mk(0., 0.);
#pragma unroll
        for (int k = 0; k < NS; ++k) zx = add(zx, dvs(x[k], n));
        const double third = 1. / 3.;
        double zvol = 0.;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const d2 p1 = x[k], p2 = x[nxt<NS>(k)];
            const double sa = 0.5 * cross(sub(p2, p1), sub(zx, p1));
            const double sv = third * sa * (p1.x + p2.x + zx.x);
            zvol += sv;
            if (sv <= 0.) bad += 1;
        }
        // calcWork
        double w = 0., et = a.zetot[z];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int p1 = pp[k], p2 = pp[nxt<NS>(k)];
            const d2 sftot = mk(a.sfwx[sb + k], a.sfwy[sb + k]);
            const double sd1 = dot(sftot, mk(a.vsx[p1], a.vsy[p1]));
            const double sd2 = dot(neg(sftot), mk(a.vsx[p2], a.vsy[p2]));
            const double dwork = -dth * (sd1 * a.pxpx[p1] + sd2 * a.pxpx[p2]);
            et += dwork;
            w += dwork;
        }
        const double zvol0 = a.zvol[z];
        const double zm = a.zm[z];
        // calcWorkRate, calcEnergy, calcRho (corr), calcDvol
        const double dtinv = 1. / dt;
        const double dvol = zvol - zvol0;
        a.zwrate[z] = (w + a.zp[z] * dvol) * dtinv;
        a.zetot[z] = et;
        const double fuzz = 1.e-99;
        a.ze[z] = et / (zm + fuzz);
        a.zr[z] = zm / zvol;
        a.zvol[z] = zvol;
        const double zdvov = fabs((zvol - zvol0) / zvol0);
        dvov = (zdvov > dvov) ? zdvov : dvov;
    }
    const double m = block_max(dvov);
    const int b = block_sum(bad);
    if (threadIdx.x == 0) {
        a.part_dv[boff + blockIdx.x] = m;
        a.part_bc[boff + blockIdx.x] = b;
    }
}

// K4: reductions over block partials and the end-of-cycle scalars.
__global__ void __launch_bounds__(PN_BLOCK) k_fused_finalize(FusedArgs a, int nparts, double cflv,
                                                             int cstop, double dtmax, double dtinit,
                                                             double dtfac, double tstop) {
    Ctl* c = a.ctl;
    if (c->done) return;
    double mn = 1.e99, mx = 1.e-99;
    int bp = 0, bc = 0;
    for (int i = threadIdx.x; i < nparts; i += PN_BLOCK) {
        mn = (a.part_dt[i] < mn) ? a.part_dt[i] : mn;
        mx = (a.part_dv[i] > mx) ? a.part_dv[i] : mx;
        bp += a.part_bp[i];
        bc += a.part_bc[i];
    }
    mn = block_min(mn);
    mx = block_max(mx);
    bp = block_sum(bp);
    bc = block_sum(bc);
    if (threadIdx.x != 0) return;
    c->dtnew = mn;
    c->dvovmax = mx;
    c->numsbad_pred = bp;
    c->numsbad_corr = bc;
    // calcDtTask, src/Hydro.cc:1647-1667
    double r = 1.e99;
    if (mn < r) r = mn;
    const double d2v = c->dt * cflv / mx;
    if (d2v < r) r = d2v;
    c->dtrec = r;
    if (bp > 0 || bc > 0) {
        c->err_cycle = c->cycle;
        c->err_pred = bp > 0 ? 1 : 0;
        c->done = 1;
        return;
    }
    c->time = c->time + c->dt;
    c->cycle = c->cycle + 1;
    c->dtlast = c->dt;
    c->done = (c->cycle >= cstop || c->time >= tstop) ? 1 : 0;
    // calcGlobalDtTask, src/Driver.cc:213-245, for the next cycle
    double dt = dtmax;
    const double dtrecover = dtfac * c->dtlast;
    if (dtrecover < dt) dt = dtrecover;
    if ((tstop - c->time) < dt) dt = tstop - c->time;
    if (c->dtrec < dt) dt = c->dtrec;
    c->dt = dt;
}

FusedArgs make_args(pennant_problem* pb) {
    FusedArgs a;
    const Params& q = pb->prm;
    a.mapsp1 = pb->dm.mapsp1; a.pcbeg = pb->dm.pcbeg; a.pcside = pb->dm.pcside;
    a.smf = pb->f("smf"); a.nump = pb->dm.nump;
    a.pxx = pb->f("px", 0); a.pxy = pb->f("px", 1); a.pux = pb->f("pu", 0); a.puy = pb->f("pu", 1);
    a.zr = pb->f("zr"); a.ze = pb->f("ze"); a.zetot = pb->f("zetot"); a.zwrate = pb->f("zwrate");
    a.zvol = pb->f("zvol"); a.zp = pb->f("zp"); a.zm = pb->f("zm");
    a.cfx = pb->cfx; a.cfy = pb->cfy; a.cmass = pb->cmass; a.sfwx = pb->sfwx; a.sfwy = pb->sfwy;
    a.vsx = pb->vsx; a.vsy = pb->vsy; a.pxpx = pb->pxpx; a.pbc = pb->pbc;
    a.part_dt = pb->part_d; a.part_dv = pb->part_d + pb->nparts;
    a.part_bp = pb->part_i; a.part_bc = pb->part_i + pb->nparts;
    a.ctl = pb->ctl;
    a.gamma = q.gamma; a.ssmin = q.ssmin; a.alfa = q.alfa; a.qgamma = q.qgamma;
    a.q1 = q.q1; a.q2 = q.q2; a.cfl = q.cfl;
    a.nbcs = pb->nbcs;
    for (int i = 0; i < 8; ++i) { a.bcvx[i] = pb->bcvx[i]; a.bcvy[i] = pb->bcvy[i]; }
    return a;
}

template <int NS>
void launch_pred(const FusedArgs& a, int zb, int ze, int sb, int boff, hipStream_t s) {
    hipLaunchKernelGGL(k_fused_zone_pred<NS>, dim3(nblocks(ze - zb)), dim3(PN_BLOCK), 0, s, a, zb,
                       ze, sb, boff);
}
template <int NS>
void launch_corr(const FusedArgs& a, int zb, int ze, int sb, int boff, hipStream_t s) {
    hipLaunchKernelGGL(k_fused_zone_corr<NS>, dim3(nblocks(ze - zb)), dim3(PN_BLOCK), 0, s, a, zb,
                       ze, sb, boff);
}

}  // namespace

namespace pn {

// Builds the zone runs, the per-point BC bits and the fused-path buffers.
int fused_setup(pennant_problem* pb) {
    const HostMesh& h = pb->hm;
    pb->zruns.clear();
    int nparts = 0;
    for (int z = 0; z < h.numz;) {
        int e = z + 1;
        while (e < h.numz && h.znump[e] == h.znump[z]) ++e;
        pb->zruns.push_back(z);
        pb->zruns.push_back(e);
        pb->zruns.push_back(h.znump[z]);
        nparts += nblocks(e - z);
        z = e;
    }
    pb->nparts = nparts;
    if (pb->bcs.size() > 8) return PENNANT_EINVAL;
    pb->nbcs = (int)pb->bcs.size();
    std::vector<unsigned char> bits((size_t)h.nump, 0);
    const double eps = 1.e-12;
    int bi = 0;
    auto mark = [&](double b, bool xplane) {
        for (int p = 0; p < h.nump; ++p) {
            const double c = xplane ? h.px[p] : h.py[p];
            if (std::fabs(c - b) < eps) bits[p] |= (unsigned char)(1u << bi);
        }
        pb->bcvx[bi] = xplane ? 1. : 0.;
        pb->bcvy[bi] = xplane ? 0. : 1.;
        ++bi;
    };
    for (double b : pb->prm.bcx) mark(b, true);
    for (double b : pb->prm.bcy) mark(b, false);
    auto al = [&](size_t n, size_t sz) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, n * sz > 0 ? n * sz : sz) != hipSuccess) return nullptr;
        if (hipMemset(p, 0, n * sz > 0 ? n * sz : sz) != hipSuccess) { (void)hipFree(p); return nullptr; }
        pb->allocs.push_back(p);
        return p;
    };
    const size_t S = h.nums, P = h.nump;
    pb->cfx = (double*)al(S, 8); pb->cfy = (double*)al(S, 8); pb->cmass = (double*)al(S, 8);
    pb->sfwx = (double*)al(S, 8); pb->sfwy = (double*)al(S, 8);
    pb->vsx = (double*)al(P, 8); pb->vsy = (double*)al(P, 8); pb->pxpx = (double*)al(P, 8);
    pb->pbc = (unsigned char*)al(P, 1);
    pb->part_d = (double*)al(2 * (size_t)nparts, 8);
    pb->part_i = (int*)al(2 * (size_t)nparts, 4);
    if (!pb->cfx || !pb->cfy || !pb->cmass || !pb->sfwx || !pb->sfwy || !pb->vsx || !pb->vsy ||
        !pb->pxpx || !pb->pbc || !pb->part_d || !pb->part_i)
        return (int)hipErrorOutOfMemory;
    return (int)hipMemcpy(pb->pbc, bits.data(), bits.size(), hipMemcpyHostToDevice);
}

int fused_cycle(pennant_problem* pb, hipEvent_t* ev) {
    const FusedArgs a = make_args(pb);
    const Params& q = pb->prm;
    hipStream_t s = pb->stream;
    const HostMesh& h = pb->hm;
    if (ev) (void)hipEventRecord(ev[0], s);
    int boff = 0;
    for (size_t r = 0; r < pb->zruns.size(); r += 3) {
        const int zb = pb->zruns[r], ze = pb->zruns[r + 1], ns = pb->zruns[r + 2];
        const int sb = h.zsbeg[zb];
        switch (ns) {
            case 3: launch_pred<3>(a, zb, ze, sb, boff, s); break;
            case 4: launch_pred<4>(a, zb, ze, sb, boff, s); break;
            case 5: launch_pred<5>(a, zb, ze, sb, boff, s); break;
            case 6: launch_pred<6>(a, zb, ze, sb, boff, s); break;
            case 7: launch_pred<7>(a, zb, ze, sb, boff, s); break;
            case 8: launch_pred<8>(a, zb, ze, sb, boff, s); break;
            default: return PENNANT_EINVAL;
        }
        boff += nblocks(ze - zb);
    }
    if (ev) (void)hipEventRecord(ev[1], s);
    hipLaunchKernelGGL(k_fused_point, dim3(nblocks(h.nump)), dim3(PN_BLOCK), 0, s, a);
    if (ev) (void)hipEventRecord(ev[2], s);
    boff = 0;
    for (size_t r = 0; r < pb->zruns.size(); r += 3) {
        const int zb = pb->zruns[r], ze = pb->zruns[r + 1], ns = pb->zruns[r + 2];
        const int sb = h.zsbeg[zb];
        switch (ns) {
            case 3: launch_corr<3>(a, zb, ze, sb, boff, s); break;
            case 4: launch_corr<4>(a, zb, ze, sb, boff, s); break;
            case 5: launch_corr<5>(a, zb, ze, sb, boff, s); break;
            case 6: launch_corr<6>(a, zb, ze, sb, boff, s); break;
            case 7: launch_corr<7>(a, zb, ze, sb, boff, s); break;
            case 8: launch_corr<8>(a, zb, ze, sb, boff, s); break;
            default: return PENNANT_EINVAL;
        }
        boff += nblocks(ze - zb);
    }
    if (ev) (void)hipEventRecord(ev[3], s);
    hipLaunchKernelGGL(k_fused_finalize, dim3(1), dim3(PN_BLOCK), 0, s, a, pb->nparts, q.cflv,
                       q.cstop, q.dtmax, q.dtinit, q.dtfac, q.tstop);
    if (ev) (void)hipEventRecord(ev[4], s);
    return (int)hipGetLastError();
}

}  // namespace pn
