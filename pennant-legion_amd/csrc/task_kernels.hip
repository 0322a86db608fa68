// Task entry points: one gfx950 kernel (or kernel pair) per Legion task ID on
// the PENNANT hydro cycle.  Each is the HIP counterpart of the reference's
// CUDA variant for that task ID, restated from the serial CPU task body it must
// agree with (file:line on each kernel) but laid out for CDNA:
//   * zone sums run in one thread per zone over the zone's contiguous sides in
//     ascending order, starting from the same identity as the CPU body, so the
//     result is the serial sum bit for bit (the reference's GPU variants use
//     double atomics, src/Mesh.hh:158-182, and are order-nondeterministic);
//   * point sums (corner mass, corner force) gather over the point-corner CSR
//     in ascending side order instead of scattering with atomics;
//   * min/max reductions use wave64 shuffles and per-block partials, then a
//     one-block finalize kernel (the reference folds warp-32 partials into a
//     DeferredReduction, src/CudaHelp.hh:14-41).
#include "common.h"
#include "reduce.h"

using namespace pn;

namespace {

// ---------------------------------------------------------------------------
// Mesh tasks

// calcCtrsTask, src/Mesh.cc:797-837
__global__ void __launch_bounds__(PN_BLOCK) k_calc_ctrs(pennant_mesh m, pennant_cvec px,
                                                        pennant_vec ex, pennant_vec zx) {
    const int z = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (z >= m.numz) return;
    const int sb = m.zsbeg[z], se = m.zsbeg[z + 1];
    const int n = m.znump[z];
    d2 acc = mk(0., 0.);
    for (int s = sb; s < se; ++s) {
        const d2 px1 = ld(px, m.mapsp1[s]);
        const d2 px2 = ld(px, m.mapsp2[s]);
        st(ex, s, scl(0.5, add(px1, px2)));
        acc = add(acc, dvs(px1, (double)n));
    }
    st(zx, z, acc);
}

// calcVolsTask, src/Mesh.cc:928-989; partial bad-side counts per block
__global__ void __launch_bounds__(PN_BLOCK) k_calc_vols(pennant_mesh m, pennant_cvec px,
                                                        pennant_cvec zx, double* sarea,
                                                        double* svol, double* zarea,
                                                        double* zvol, int* numsbad) {
    const int z = blockIdx.x * PN_BLOCK + threadIdx.x;
    int count = 0;
    if (z < m.numz) {
        const double third = 1. / 3.;
        const d2 c = ld(zx, z);
        double za = 0., zv = 0.;
        for (int s = m.zsbeg[z]; s < m.zsbeg[z + 1]; ++s) {
            const d2 px1 = ld(px, m.mapsp1[s]);
            const d2 px2 = ld(px, m.mapsp2[s]);
            const double sa = 0.5 * cross(sub(px2, px1), sub(c, px1));
            const double sv = third * sa * (px1.x + px2.x + c.x);
            sarea[s] = sa;
            svol[s] = sv;
            za += sa;
            zv += sv;
            if (sv <= 0.) count += 1;
        }
        zarea[z] = za;
        zvol[z] = zv;
    }
    // an integer sum is exact in any order: one atomic per block
    count = block_sum(count);
    if (threadIdx.x == 0 && count > 0) atomicAdd(numsbad, count);
}

__global__ void k_zero_int(int32_t* out) { *out = 0; }

// calcSideFracsTask, src/Mesh.cc:1085-1101
__global__ void __launch_bounds__(PN_BLOCK) k_calc_side_fracs(pennant_mesh m, const double* sarea,
                                                              const double* zarea, double* smf) {
    const int s = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (s >= m.nums) return;
    smf[s] = sarea[s] / zarea[m.mapsz[s]];
}

// calcSurfVecsTask, src/Mesh.cc:1104-1123
__global__ void __launch_bounds__(PN_BLOCK) k_calc_surf_vecs(pennant_mesh m, pennant_cvec ex,
                                                             pennant_cvec zx, pennant_vec ssurf) {
    const int s = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (s >= m.nums) return;
    st(ssurf, s, rot_ccw(sub(ld(ex, s), ld(zx, m.mapsz[s]))));
}

// calcEdgeLenTask, src/Mesh.cc:1151-1179
__global__ void __launch_bounds__(PN_BLOCK) k_calc_edge_len(pennant_mesh m, pennant_cvec px,
                                                            double* elen) {
    const int s = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (s >= m.nums) return;
    elen[s] = len(sub(ld(px, m.mapsp2[s]), ld(px, m.mapsp1[s])));
}

// calcCharLenTask, src/Mesh.cc:1216-1244
__global__ void __launch_bounds__(PN_BLOCK) k_calc_char_len(pennant_mesh m, const double* sarea,
                                                            const double* elen, double* zdl) {
    const int z = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (z >= m.numz) return;
    const double fac = (m.znump[z] == 3 ? 3. : 4.);
    double dl = 1.e99;
    for (int s = m.zsbeg[z]; s < m.zsbeg[z + 1]; ++s) {
        const double sdl = fac * sarea[s] / elen[s];
        dl = smin(dl, sdl);
    }
    zdl[z] = dl;
}

// ---------------------------------------------------------------------------
// Hydro tasks

// advPosHalfTask, src/Hydro.cc:986-1006
__global__ void __launch_bounds__(PN_BLOCK) k_adv_pos_half(int nump, pennant_cvec px0,
                                                           pennant_cvec pu0, pennant_vec pxp,
                                                           const double* dtp) {
    const int p = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (p >= nump) return;
    const double dth = 0.5 * (*dtp);
    st(pxp, p, add(ld(px0, p), scl(dth, ld(pu0, p))));
}

// calcRhoTask, src/Hydro.cc:1034-1054
__global__ void __launch_bounds__(PN_BLOCK) k_calc_rho(int numz, const double* zm,
                                                       const double* zvol, double* zr) {
    const int z = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (z >= numz) return;
    zr[z] = zm[z] / zvol[z];
}

// calcCrnrMassTask, src/Hydro.cc:1083-1117, as a gather over the point's
// corners in ascending side order
__global__ void __launch_bounds__(PN_BLOCK) k_calc_crnr_mass(pennant_mesh m, const double* smf,
                                                             const double* zr,
                                                             const double* zarea,
                                                             double* pmaswt) {
    const int p = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (p >= m.nump) return;
    double acc = pmaswt[p];
    for (int k = m.pcbeg[p]; k < m.pcbeg[p + 1]; ++k) {
        const int s = m.pcside[k];
        const int z = m.mapsz[s];
        const double mwt = zr[z] * zarea[z] * 0.5 * (smf[s] + smf[m.mapss3[s]]);
        acc += mwt;
    }
    pmaswt[p] = acc;
}

// sumCrnrForceTask, src/Hydro.cc:1159-1193, as a gather
__global__ void __launch_bounds__(PN_BLOCK) k_sum_crnr_force(pennant_mesh m, pennant_cvec sfp,
                                                             pennant_cvec sfq, pennant_cvec sft,
                                                             pennant_vec pf) {
    const int p = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (p >= m.nump) return;
    d2 acc = ld(pf, p);
    for (int k = m.pcbeg[p]; k < m.pcbeg[p + 1]; ++k) {
        const int s = m.pcside[k];
        const int s3 = m.mapss3[s];
        const d2 f = add(add(ld(sfp, s), ld(sfq, s)), ld(sft, s));
        const d2 f3 = add(add(ld(sfp, s3), ld(sfq, s3)), ld(sft, s3));
        acc = add(acc, sub(f, f3));
    }
    st(pf, p, acc);
}

// calcAccelTask, src/Hydro.cc:1235-1254
__global__ void __launch_bounds__(PN_BLOCK) k_calc_accel(int nump, pennant_cvec pf,
                                                         const double* pmaswt, pennant_vec pa) {
    const int p = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (p >= nump) return;
    const double fuzz = 1.e-99;
    st(pa, p, dvs(ld(pf, p), smax(pmaswt[p], fuzz)));
}

// advPosFullTask, src/Hydro.cc:1281-1306
__global__ void __launch_bounds__(PN_BLOCK) k_adv_pos_full(int nump, pennant_cvec px0,
                                                           pennant_cvec pu0, pennant_cvec pa,
                                                           pennant_vec px, pennant_vec pu,
                                                           const double* dtp) {
    const int p = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (p >= nump) return;
    const double dt = *dtp;
    const d2 u0 = ld(pu0, p);
    const d2 u = add(u0, scl(dt, ld(pa, p)));
    st(pu, p, u);
    st(px, p, add(ld(px0, p), scl(dt * 0.5, add(u0, u))));
}

// calcWorkTask, src/Hydro.cc:1339-1406
__global__ void __launch_bounds__(PN_BLOCK) k_calc_work(pennant_mesh m, pennant_cvec sfp,
                                                        pennant_cvec sfq, pennant_cvec pu,
                                                        pennant_cvec pu0, pennant_cvec pxp,
                                                        double* zw, double* zetot,
                                                        const double* dtp) {
    const int z = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (z >= m.numz) return;
    const double dth = 0.5 * (*dtp);
    double w = 0., et = zetot[z];
    for (int s = m.zsbeg[z]; s < m.zsbeg[z + 1]; ++s) {
        const int p1 = m.mapsp1[s], p2 = m.mapsp2[s];
        const d2 sftot = add(ld(sfp, s), ld(sfq, s));
        const double sd1 = dot(sftot, add(ld(pu0, p1), ld(pu, p1)));
        const double sd2 = dot(neg(sftot), add(ld(pu0, p2), ld(pu, p2)));
        const double dwork = -dth * (sd1 * pxp.x[p1] + sd2 * pxp.x[p2]);
        et += dwork;
        w += dwork;
    }
    zetot[z] = et;
    zw[z] = w;
}

// calcWorkRateTask, src/Hydro.cc:1483-1510
__global__ void __launch_bounds__(PN_BLOCK) k_calc_work_rate(int numz, const double* zvol0,
                                                             const double* zvol, const double* zw,
                                                             const double* zp, double* zwrate,
                                                             const double* dtp) {
    const int z = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (z >= numz) return;
    const double dtinv = 1. / (*dtp);
    const double dvol = zvol[z] - zvol0[z];
    zwrate[z] = (zw[z] + zp[z] * dvol) * dtinv;
}

// calcEnergyTask, src/Hydro.cc:1545-1564
__global__ void __launch_bounds__(PN_BLOCK) k_calc_energy(int numz, const double* zetot,
                                                          const double* zm, double* ze) {
    const int z = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (z >= numz) return;
    const double fuzz = 1.e-99;
    ze[z] = zetot[z] / (zm[z] + fuzz);
}

// calcDtTask, src/Hydro.cc:1647-1667
__global__ void k_calc_dt(const double* dtlast, const double* dtnew, const double* dvovmax,
                          double cflv, double* dtrec) {
    double r = 1.e99;
    const double a = *dtnew;
    if (a < r) r = a;
    const double b = *dtlast * cflv / *dvovmax;
    if (b < r) r = b;
    *dtrec = r;
}

// calcGlobalDtTask, src/Driver.cc:213-245
__global__ void k_calc_global_dt(const double* dtlast, const double* dtrec, const double* time,
                                 int cycle, double dtmax, double dtinit, double dtfac,
                                 double tstop, double* dtout) {
    double dt = dtmax;
    if (cycle == 0) {
        if (dtinit < dt) dt = dtinit;
    } else {
        const double dtrecover = dtfac * (*dtlast);
        if (dtrecover < dt) dt = dtrecover;
    }
    if ((tstop - *time) < dt) dt = tstop - *time;
    if (cycle > 0) {
        if (*dtrec < dt) dt = *dtrec;
    }
    *dtout = dt;
}

// ---------------------------------------------------------------------------
// PolyGas, TTS

// PolyGas::calcStateHalfTask, src/PolyGas.cc:63-108
__global__ void __launch_bounds__(PN_BLOCK) k_calc_state_half(int numz, const double* zr,
                                                              const double* zvolp,
                                                              const double* zvol0,
                                                              const double* ze,
                                                              const double* zwrate,
                                                              const double* zm, double* zp,
                                                              double* zss, double gamma,
                                                              double ssmin, const double* dtp) {
    const int z = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (z >= numz) return;
    const double dth = 0.5 * (*dtp);
    const double gm1 = gamma - 1.;
    const double ssmin2 = smax(ssmin * ssmin, 1.e-99);
    const double r = zr[z];
    const double e = smax(ze[z], 0.);
    const double p = gm1 * r * e;
    const double pre = gm1 * e;
    const double per = gm1 * r;
    const double csqd = smax(ssmin2, pre + per * p / (r * r));
    const double ss = __builtin_sqrt(csqd);
    const double minv = 1. / zm[z];
    const double dv = (zvolp[z] - zvol0[z]) * minv;
    const double bulk = r * csqd;
    const double denom = 1. + 0.5 * per * dv;
    const double src = zwrate[z] * dth * minv;
    zp[z] = p + (per * src - r * bulk * dv) / denom;
    zss[z] = ss;
}

// PolyGas::calcForceTask, src/PolyGas.cc:164-185
__global__ void __launch_bounds__(PN_BLOCK) k_calc_force_pgas(pennant_mesh m, pennant_cvec ssurf,
                                                              const double* zp, pennant_vec sf) {
    const int s = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (s >= m.nums) return;
    const double p = zp[m.mapsz[s]];
    st(sf, s, scl(-p, ld(ssurf, s)));
}

// TTS::calcForceTask, src/TTS.cc:55-104
__global__ void __launch_bounds__(PN_BLOCK) k_calc_force_tts(pennant_mesh m, const double* sarea,
                                                             const double* smf, pennant_cvec ssurf,
                                                             const double* zarea, const double* zr,
                                                             const double* zss, pennant_vec sf,
                                                             double alfa, double ssmin) {
    const int s = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (s >= m.nums) return;
    const int z = m.mapsz[s];
    const double vfacinv = zarea[z] / sarea[s];
    const double r = zr[z];
    const double srho = r * smf[s] * vfacinv;
    double sstmp = smax(zss[z], ssmin);
    sstmp = alfa * sstmp * sstmp;
    const double sdp = sstmp * (srho - r);
    st(sf, s, scl(-sdp, ld(ssurf, s)));
}

// ---------------------------------------------------------------------------
// QCS

// QCS::setCornerDivTask, src/QCS.cc:116-242 (one thread per zone: zone
// velocity first, then the zone's corners)
__global__ void __launch_bounds__(PN_BLOCK) k_set_corner_div(pennant_mesh m, pennant_cvec ex,
                                                             const double* elen, pennant_cvec zx,
                                                             pennant_cvec px, pennant_cvec pu,
                                                             pennant_vec zuc, double* carea,
                                                             double* ccos, double* cdiv,
                                                             double* cevol, double* cdu) {
    const int z = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (z >= m.numz) return;
    const int sb = m.zsbeg[z], se = m.zsbeg[z + 1];
    const int n = m.znump[z];
    d2 uc = mk(0., 0.);
    for (int s = sb; s < se; ++s) uc = add(uc, dvs(ld(pu, m.mapsp1[s]), (double)n));
    st(zuc, z, uc);
    const d2 xp2 = ld(zx, z);
    const d2 up2 = uc;
    for (int c = sb; c < se; ++c) {
        const int s2 = c;
        const int s = m.mapss3[s2];
        const int p = m.mapsp2[s];
        const int p1 = m.mapsp1[s];
        const int p2 = m.mapsp2[s2];
        const d2 up0 = ld(pu, p);
        const d2 xp0 = ld(px, p);
        const d2 up1 = scl(0.5, add(up0, ld(pu, p2)));
        const d2 xp1 = ld(ex, s2);
        const d2 up3 = scl(0.5, add(ld(pu, p1), up0));
        const d2 xp3 = ld(ex, s);
        const double cvolume = 0.5 * cross(sub(xp2, xp0), sub(xp3, xp1));
        carea[c] = cvolume;
        const d2 v1 = sub(xp3, xp0);
        const d2 v2 = sub(xp1, xp0);
        const double de1 = elen[s];
        const double de2 = elen[s2];
        const double minelen = smin(de1, de2);
        ccos[c] = ((minelen < 1.e-12) ? 0. : 4. * dot(v1, v2) / (de1 * de2));
        const double div = (cross(sub(up2, up0), sub(xp3, xp1)) - cross(sub(up3, up1), sub(xp2, xp0))) /
                           (2.0 * cvolume);
        cdiv[c] = div;
        const d2 dxx1 = scl(0.5, sub(sub(add(xp1, xp2), xp0), xp3));
        const d2 dxx2 = scl(0.5, sub(sub(add(xp2, xp3), xp0), xp1));
        const double dx1 = len(dxx1);
        const double dx2 = len(dxx2);
        const d2 duav = scl(0.25, add(add(add(up0, up1), up2), up3));
        const double test1 = fabs(dot(dxx1, duav) * dx2);
        const double test2 = fabs(dot(dxx2, duav) * dx1);
        const double num = (test1 > test2 ? dx1 : dx2);
        const double den = (test1 > test2 ? dx2 : dx1);
        const double r = num / den;
        double evol = __builtin_sqrt(4.0 * cvolume * r);
        evol = smin(evol, 2.0 * minelen);
        const double dv1 = len2(sub(sub(add(up1, up2), up0), up3));
        const double dv2 = len2(sub(sub(add(up2, up3), up0), up1));
        double du = __builtin_sqrt(smax(dv1, dv2));
        evol = (div < 0.0 ? evol : 0.);
        du = (div < 0.0 ? du : 0.);
        cevol[c] = evol;
        cdu[c] = du;
    }
}

// QCS::setQCnForceTask, src/QCS.cc:246-327
__global__ void __launch_bounds__(PN_BLOCK) k_set_qcn_force(pennant_mesh m, const double* elen,
                                                            const double* cdiv, const double* cdu,
                                                            const double* cevol, const double* zrp,
                                                            const double* zss, pennant_cvec pu,
                                                            double* crmu, pennant_vec cqe1,
                                                            pennant_vec cqe2, double qgamma,
                                                            double q1, double q2) {
    const int c = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (c >= m.nums) return;
    const double gammap1 = qgamma + 1.0;
    const int z = m.mapsz[c];
    const double ztmp2 = q2 * 0.25 * gammap1 * cdu[c];
    const double ztmp1 = q1 * zss[z];
    const double zkur = ztmp2 + __builtin_sqrt(ztmp2 * ztmp2 + ztmp1 * ztmp1);
    const double mu = zkur * zrp[z] * cevol[c];
    const double cm = ((cdiv[c] > 0.0) ? 0. : mu);
    crmu[c] = cm;
    const int s2 = c;
    const int s = m.mapss3[s2];
    const int p = m.mapsp2[s];
    const int p1 = m.mapsp1[s];
    const int p2 = m.mapsp2[s2];
    const d2 u = ld(pu, p);
    st(cqe1, c, dvs(scl(cm, sub(u, ld(pu, p1))), elen[s]));
    st(cqe2, c, dvs(scl(cm, sub(ld(pu, p2), u)), elen[s2]));
}

// QCS::setForceTask, src/QCS.cc:331-380 (one thread per zone so that the
// clamped ccos/cw of the neighbouring corner c2 = ss4[s] are ready)
__global__ void __launch_bounds__(PN_BLOCK) k_set_force_qcs(pennant_mesh m, const double* carea,
                                                            pennant_cvec cqe1, pennant_cvec cqe2,
                                                            const double* elen, double* ccos,
                                                            double* cw, pennant_vec sfq) {
    const int z = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (z >= m.numz) return;
    const int sb = m.zsbeg[z], se = m.zsbeg[z + 1];
    for (int c = sb; c < se; ++c) {
        const double cc = ccos[c];
        const double csin2 = 1.0 - cc * cc;
        cw[c] = ((csin2 < 1.e-4) ? 0. : carea[c] / csin2);
        ccos[c] = ((csin2 < 1.e-4) ? 0. : cc);
    }
    for (int s = sb; s < se; ++s) {
        const int c1 = s, c2 = m.mapss4[s];
        const double el = elen[s];
        const double cw1 = cw[c1], cw2 = cw[c2];
        const double ccos1 = ccos[c1], ccos2 = ccos[c2];
        const d2 cqe11 = ld(cqe1, c1), cqe12 = ld(cqe1, c2);
        const d2 cqe21 = ld(cqe2, c1), cqe22 = ld(cqe2, c2);
        const d2 f = add(scl(cw1, add(cqe21, scl(ccos1, cqe11))),
                         scl(cw2, add(cqe12, scl(ccos2, cqe22))));
        st(sfq, s, dvs(f, el));
    }
}

// QCS::setVelDiffTask, src/QCS.cc:384-448
__global__ void __launch_bounds__(PN_BLOCK) k_set_vel_diff(pennant_mesh m, const double* elen,
                                                           const double* zss, pennant_cvec px,
                                                           pennant_cvec pu, double* ztmp,
                                                           double* zdu, double q1, double q2) {
    const int z = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (z >= m.numz) return;
    double t = 0.;
    for (int s = m.zsbeg[z]; s < m.zsbeg[z + 1]; ++s) {
        const int p1 = m.mapsp1[s], p2 = m.mapsp2[s];
        const d2 dx = sub(ld(px, p2), ld(px, p1));
        const d2 du = sub(ld(pu, p2), ld(pu, p1));
        const double lenx = elen[s];
        double dux = dot(du, dx);
        dux = (lenx > 0. ? fabs(dux) / lenx : 0.);
        t = smax(t, dux);
    }
    ztmp[z] = t;
    zdu[z] = q1 * zss[z] + 2. * q2 * t;
}

// ---------------------------------------------------------------------------
// HydroBC::applyFixedBCTask, src/HydroBC.cc:150-183
__global__ void __launch_bounds__(PN_BLOCK) k_apply_fixed_bc(const int32_t* mapbp, int numb,
                                                             double vx, double vy, pennant_vec pf,
                                                             pennant_vec pu0) {
    const int b = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (b >= numb) return;
    const int p = mapbp[b];
    const d2 u = mk(vx, vy);
    st(pu0, p, project(ld(pu0, p), u));
    st(pf, p, project(ld(pf, p), u));
}

inline int launch_status() { return (int)hipGetLastError(); }

}  // namespace

// ===========================================================================
// C ABI

extern "C" {

size_t pennant_hip_reduce_workspace_bytes(void) {
    return (size_t)PN_RED_BLOCKS * sizeof(double);
}

int pennant_hip_calc_ctrs(const pennant_mesh* m, pennant_cvec px, pennant_vec ex,
                          pennant_vec zx, hipStream_t st) {
    if (m->numz > 0)
        hipLaunchKernelGGL(k_calc_ctrs, dim3(nblocks(m->numz)), dim3(PN_BLOCK), 0, st, *m, px, ex, zx);
    return launch_status();
}

int pennant_hip_calc_vols(const pennant_mesh* m, pennant_cvec px, pennant_cvec zx,
                          double* sarea, double* svol, double* zarea, double* zvol,
                          int32_t* numsbad, void* ws, hipStream_t st) {
    (void)ws;
    if (!numsbad) return PENNANT_EINVAL;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, numsbad);
    if (m->numz > 0)
        hipLaunchKernelGGL(k_calc_vols, dim3(nblocks(m->numz)), dim3(PN_BLOCK), 0, st, *m, px, zx,
                           sarea, svol, zarea, zvol, numsbad);
    return launch_status();
}

int pennant_hip_calc_side_fracs(const pennant_mesh* m, const double* sarea, const double* zarea,
                                double* smf, hipStream_t st) {
    if (m->nums > 0)
        hipLaunchKernelGGL(k_calc_side_fracs, dim3(nblocks(m->nums)), dim3(PN_BLOCK), 0, st, *m,
                           sarea, zarea, smf);
    return launch_status();
}

int pennant_hip_calc_surf_vecs(const pennant_mesh* m, pennant_cvec ex, pennant_cvec zx,
                               pennant_vec ssurf, hipStream_t st) {
    if (m->nums > 0)
        hipLaunchKernelGGL(k_calc_surf_vecs, dim3(nblocks(m->nums)), dim3(PN_BLOCK), 0, st, *m, ex,
                           zx, ssurf);
    return launch_status();
}

int pennant_hip_calc_edge_len(const pennant_mesh* m, pennant_cvec px, double* elen,
                              hipStream_t st) {
    if (m->nums > 0)
        hipLaunchKernelGGL(k_calc_edge_len, dim3(nblocks(m->nums)), dim3(PN_BLOCK), 0, st, *m, px,
                           elen);
    return launch_status();
}

int pennant_hip_calc_char_len(const pennant_mesh* m, const double* sarea, const double* elen,
                              double* zdl, hipStream_t st) {
    if (m->numz > 0)
        hipLaunchKernelGGL(k_calc_char_len, dim3(nblocks(m->numz)), dim3(PN_BLOCK), 0, st, *m,
                           sarea, elen, zdl);
    return launch_status();
}

int pennant_hip_adv_pos_half(const pennant_mesh* m, pennant_cvec px0, pennant_cvec pu0,
                             pennant_vec pxp, const double* dt, hipStream_t st) {
    if (m->nump > 0)
        hipLaunchKernelGGL(k_adv_pos_half, dim3(nblocks(m->nump)), dim3(PN_BLOCK), 0, st, m->nump,
                           px0, pu0, pxp, dt);
    return launch_status();
}

int pennant_hip_calc_rho(const pennant_mesh* m, const double* zm, const double* zvol, double* zr,
                         hipStream_t st) {
    if (m->numz > 0)
        hipLaunchKernelGGL(k_calc_rho, dim3(nblocks(m->numz)), dim3(PN_BLOCK), 0, st, m->numz, zm,
                           zvol, zr);
    return launch_status();
}

int pennant_hip_calc_crnr_mass(const pennant_mesh* m, const double* smf, const double* zr,
                               const double* zarea, double* pmaswt, hipStream_t st) {
    if (m->nump > 0)
        hipLaunchKernelGGL(k_calc_crnr_mass, dim3(nblocks(m->nump)), dim3(PN_BLOCK), 0, st, *m,
                           smf, zr, zarea, pmaswt);
    return launch_status();
}

int pennant_hip_sum_crnr_force(const pennant_mesh* m, pennant_cvec sfp, pennant_cvec sfq,
                               pennant_cvec sft, pennant_vec pf, hipStream_t st) {
    if (m->nump > 0)
        hipLaunchKernelGGL(k_sum_crnr_force, dim3(nblocks(m->nump)), dim3(PN_BLOCK), 0, st, *m,
                           sfp, sfq, sft, pf);
    return launch_status();
}

int pennant_hip_calc_accel(const pennant_mesh* m, pennant_cvec pf, const double* pmaswt,
                           pennant_vec pa, hipStream_t st) {
    if (m->nump > 0)
        hipLaunchKernelGGL(k_calc_accel, dim3(nblocks(m->nump)), dim3(PN_BLOCK), 0, st, m->nump,
                           pf, pmaswt, pa);
    return launch_status();
}

int pennant_hip_adv_pos_full(const pennant_mesh* m, pennant_cvec px0, pennant_cvec pu0,
                             pennant_cvec pa, pennant_vec px, pennant_vec pu, const double* dt,
                             hipStream_t st) {
    if (m->nump > 0)
        hipLaunchKernelGGL(k_adv_pos_full, dim3(nblocks(m->nump)), dim3(PN_BLOCK), 0, st, m->nump,
                           px0, pu0, pa, px, pu, dt);
    return launch_status();
}

int pennant_hip_calc_work(const pennant_mesh* m, pennant_cvec sfp, pennant_cvec sfq,
                          pennant_cvec pu, pennant_cvec pu0, pennant_cvec pxp, double* zw,
                          double* zetot, const double* dt, hipStream_t st) {
    if (m->numz > 0)
        hipLaunchKernelGGL(k_calc_work, dim3(nblocks(m->numz)), dim3(PN_BLOCK), 0, st, *m, sfp,
                           sfq, pu, pu0, pxp, zw, zetot, dt);
    return launch_status();
}

int pennant_hip_calc_work_rate(const pennant_mesh* m, const double* zvol0, const double* zvol,
                               const double* zw, const double* zp, double* zwrate,
                               const double* dt, hipStream_t st) {
    if (m->numz > 0)
        hipLaunchKernelGGL(k_calc_work_rate, dim3(nblocks(m->numz)), dim3(PN_BLOCK), 0, st,
                           m->numz, zvol0, zvol, zw, zp, zwrate, dt);
    return launch_status();
}

int pennant_hip_calc_energy(const pennant_mesh* m, const double* zetot, const double* zm,
                            double* ze, hipStream_t st) {
    if (m->numz > 0)
        hipLaunchKernelGGL(k_calc_energy, dim3(nblocks(m->numz)), dim3(PN_BLOCK), 0, st, m->numz,
                           zetot, zm, ze);
    return launch_status();
}

int pennant_hip_calc_dt_new(const pennant_mesh* m, const double* zdl, const double* zdu,
                            const double* zss, double cfl, double* dtnew, void* ws,
                            hipStream_t st) {
    if (!ws || !dtnew) return PENNANT_EINVAL;
    const int nb = nred_blocks(m->numz);
    double* part = static_cast<double*>(ws);
    hipLaunchKernelGGL(k_calc_dt_new, dim3(nb), dim3(PN_BLOCK), 0, st, m->numz, zdl, zdu, zss,
                       cfl, part);
    hipLaunchKernelGGL(k_min_parts, dim3(1), dim3(PN_BLOCK), 0, st, part, nb, dtnew);
    return launch_status();
}

int pennant_hip_calc_dvol(const pennant_mesh* m, const double* zvol, const double* zvol0,
                          double* dvovmax, void* ws, hipStream_t st) {
    if (!ws || !dvovmax) return PENNANT_EINVAL;
    const int nb = nred_blocks(m->numz);
    double* part = static_cast<double*>(ws);
    hipLaunchKernelGGL(k_calc_dvol, dim3(nb), dim3(PN_BLOCK), 0, st, m->numz, zvol, zvol0, part);
    hipLaunchKernelGGL(k_max_parts, dim3(1), dim3(PN_BLOCK), 0, st, part, nb, dvovmax);
    return launch_status();
}

int pennant_hip_calc_dt(const double* dtlast, const double* dtnew, const double* dvovmax,
                        double cflv, double* dtrec, hipStream_t st) {
    hipLaunchKernelGGL(k_calc_dt, dim3(1), dim3(1), 0, st, dtlast, dtnew, dvovmax, cflv, dtrec);
    return launch_status();
}

int pennant_hip_calc_global_dt(const double* dtlast, const double* dtrec, const double* time,
                               int32_t cycle, double dtmax, double dtinit, double dtfac,
                               double tstop, double* dt, hipStream_t st) {
    hipLaunchKernelGGL(k_calc_global_dt, dim3(1), dim3(1), 0, st, dtlast, dtrec, time, cycle,
                       dtmax, dtinit, dtfac, tstop, dt);
    return launch_status();
}

int pennant_hip_calc_state_half(const pennant_mesh* m, const double* zr, const double* zvolp,
                                const double* zvol0, const double* ze, const double* zwrate,
                                const double* zm, double* zp, double* zss, double gamma,
                                double ssmin, const double* dt, hipStream_t st) {
    if (m->numz > 0)
        hipLaunchKernelGGL(k_calc_state_half, dim3(nblocks(m->numz)), dim3(PN_BLOCK), 0, st,
                           m->numz, zr, zvolp, zvol0, ze, zwrate, zm, zp, zss, gamma, ssmin, dt);
    return launch_status();
}

int pennant_hip_calc_force_pgas(const pennant_mesh* m, pennant_cvec ssurf, const double* zp,
                                pennant_vec sf, hipStream_t st) {
    if (m->nums > 0)
        hipLaunchKernelGGL(k_calc_force_pgas, dim3(nblocks(m->nums)), dim3(PN_BLOCK), 0, st, *m,
                           ssurf, zp, sf);
    return launch_status();
}

int pennant_hip_calc_force_tts(const pennant_mesh* m, const double* sarea, const double* smf,
                               pennant_cvec ssurf, const double* zarea, const double* zr,
                               const double* zss, pennant_vec sf, double alfa, double ssmin,
                               hipStream_t st) {
    if (m->nums > 0)
        hipLaunchKernelGGL(k_calc_force_tts, dim3(nblocks(m->nums)), dim3(PN_BLOCK), 0, st, *m,
                           sarea, smf, ssurf, zarea, zr, zss, sf, alfa, ssmin);
    return launch_status();
}

int pennant_hip_set_corner_div(const pennant_mesh* m, pennant_cvec ex, const double* elen,
                               pennant_cvec zx, pennant_cvec px, pennant_cvec pu,
                               pennant_vec zuc, double* carea, double* ccos, double* cdiv,
                               double* cevol, double* cdu, hipStream_t st) {
    if (m->numz > 0)
        hipLaunchKernelGGL(k_set_corner_div, dim3(nblocks(m->numz)), dim3(PN_BLOCK), 0, st, *m, ex,
                           elen, zx, px, pu, zuc, carea, ccos, cdiv, cevol, cdu);
    return launch_status();
}

int pennant_hip_set_qcn_force(const pennant_mesh* m, const double* elen, const double* cdiv,
                              const double* cdu, const double* cevol, const double* zrp,
                              const double* zss, pennant_cvec pu, double* crmu,
                              pennant_vec cqe1, pennant_vec cqe2, double qgamma, double q1,
                              double q2, hipStream_t st) {
    if (m->nums > 0)
        hipLaunchKernelGGL(k_set_qcn_force, dim3(nblocks(m->nums)), dim3(PN_BLOCK), 0, st, *m,
                           elen, cdiv, cdu, cevol, zrp, zss, pu, crmu, cqe1, cqe2, qgamma, q1, q2);
    return launch_status();
}

int pennant_hip_set_force_qcs(const pennant_mesh* m, const double* carea, pennant_cvec cqe1,
                              pennant_cvec cqe2, const double* elen, double* ccos, double* cw,
                              pennant_vec sfq, hipStream_t st) {
    if (m->numz > 0)
        hipLaunchKernelGGL(k_set_force_qcs, dim3(nblocks(m->numz)), dim3(PN_BLOCK), 0, st, *m,
                           carea, cqe1, cqe2, elen, ccos, cw, sfq);
    return launch_status();
}

int pennant_hip_set_vel_diff(const pennant_mesh* m, const double* elen, const double* zss,
                             pennant_cvec px, pennant_cvec pu, double* ztmp, double* zdu,
                             double q1, double q2, hipStream_t st) {
    if (m->numz > 0)
        hipLaunchKernelGGL(k_set_vel_diff, dim3(nblocks(m->numz)), dim3(PN_BLOCK), 0, st, *m, elen,
                           zss, px, pu, ztmp, zdu, q1, q2);
    return launch_status();
}

int pennant_hip_apply_fixed_bc(const int32_t* mapbp, int32_t numb, double vfix_x, double vfix_y,
                               pennant_vec pf, pennant_vec pu0, hipStream_t st) {
    if (numb > 0)
        hipLaunchKernelGGL(k_apply_fixed_bc, dim3(nblocks(numb)), dim3(PN_BLOCK), 0, st, mapbp,
                           numb, vfix_x, vfix_y, pf, pu0);
    return launch_status();
}

}  // extern "C"
