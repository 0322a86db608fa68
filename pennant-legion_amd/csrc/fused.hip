// Fused PENNANT cycle for gfx950.
//
// The reference issues ~36 launches per cycle (src/Hydro.cc:342-983), each a
// streaming pass over sides, zones or points with intermediate arrays in
// memory.  Every task from advPosHalf to setVelDiff only couples a zone to
// its own sides and corners, so the whole predictor half of doCycle is one
// zone-local computation; the only cross-zone couplings are the two point
// sums (corner mass, corner force) and the corrector's point positions.  The
// fused cycle is therefore three streaming passes and a one-block finalize:
//
//   K1 zone pass   advPosHalf (per corner), calcCtrs, calcVols, calcSurfVecs,
//                  calcEdgeLen, calcCharLen, calcRho, calcCrnrMass (per-corner
//                  mass), calcStateHalf, calcForcePGas, calcForceTTS,
//                  setCornerDiv, setQCnForce, setForceQCS, setVelDiff, the
//                  corner force of sumCrnrForce, calcDtNew's per-zone dt
//   K2 point pass  corner-mass and corner-force sums (ascending side order;
//                  K1 stores each corner's values at its point's slot for
//                  that corner, corner-major so that neighbouring points'
//                  slots are adjacent), applyFixedBC, calcAccel, advPosFull
//   K3 zone pass   calcCtrs, calcVols (corrector), calcWork, calcWorkRate,
//                  calcEnergy, calcRho, calcDvol's per-zone ratio
//   K4 finalize    dt minimum, dvol maximum and bad-side counts over the
//                  block partials of K1 and K3, calcDt, checkBadSides, the
//                  time/cycle update and calcGlobalDt for the next cycle
//
// Every expression is evaluated exactly as in the task kernels (and the
// reference's serial bodies), so the end-of-cycle state is bit-identical to
// the task-by-task path.  Each kernel starts by reading the device `done`
// flag, which plays the role of the reference's cycle predicate
// (src/Driver.cc:140-144): launches after the stop condition do nothing, so
// the host can queue many cycles without reading anything back.
//
// K1 has two layouts, chosen per problem (pennant_set_tuning):
//   zone-per-thread  one thread holds a whole zone in registers: no
//                    redundant zone-level arithmetic (EOS, density, dt);
//   corner-per-lane  a zone with NS sides occupies NS adjacent lanes of a
//                    64-lane wave (16 quads or 21 triangles per wave); side
//                    arrays are read and written fully coalesced and zone
//                    sums read the group's values with cross-lane shuffles in
//                    ascending corner order, so every lane of the group forms
//                    the same sum in the reference's order.
// K1 and K3 are grid-stride kernels with a bounded grid, each block writing
// one partial per reduction (no same-address atomics on the hot path).  The
// host launches K1 and K3 once per run of consecutive zones with equal side
// count (rect: one run of quads; pie: triangles, then quads).
#include <cstring>

#include "closure.h"
#include "common.h"
#include "problem.h"

using namespace pn;

#define PN_FIN_BLOCK 256     // finalize block size

namespace {

struct FusedArgs {
    // mesh
    const int32_t* mapsp1;
    const int32_t* pcbeg;
    const int32_t* pcpos;      // side -> slot of its corner record (corner_slot)
    const double* smf;
    int nump;
    int npown;                 // K2 updates points [0, npown)
    int la_npred;              // pipelined cycle: K2's last block folds a redone predictor's partials
    // state (in place)
    double* pxx; double* pxy; double* pux; double* puy;
    double* zr; double* ze; double* zetot; double* zwrate; double* zvol; double* zp;
    const double* zm;
    double* zpn;               // pipelined mixed cycle: the look-ahead predictor's zp
    // per-side intermediates
    double* cfx; double* cfy; double* cmass; double* sfwx; double* sfwy;
    // the mixed-precision form's FP32 counterparts, and smf in FP32
    float* fcrn;               // per corner slot {fx, fy, mass} (corner_slot order)
    float* fsfw;               // per side {x, y}
    const float* fsmf;
    // per-point intermediates
    double* vs2; double* pxpx;   // vs2: per point {x, y} (see point_update)
    const unsigned char* pbc;
    const int* heavy;
    int nheavy;
    // block partials
    double* part_dt; double* part_dv; int* part_bp; int* part_bc;
    Ctl* ctl;
    // physics parameters
    double gamma, ssmin, alfa, qgamma, q1, q2, cfl;
    int nbcs;
    const double* bcv;         // BC plane normals: x components [0, 8), y components [8, 16)
    int xcd;   // XCD-aware unit mapping in the grid-stride kernels
    // learned closure and training-tuple capture
    const ClosurePack* pack;
    float* cap_buf;
    long long cap;
    unsigned long long* cap_count;
    unsigned long long* cl_stats;
    // structured rect mesh (rect_nzx > 0): points and corner slots of a zone
    // from its index instead of mapsp1 and pcpos (rect_zone)
    int rect_nzx;
    unsigned long long rect_m;  // z / nzx = (z * rect_m) >> rect_sh
    int rect_sh;
    int rect_so, rect_po;       // the mesh's first side and point in the arrays (a batch's arena)
    int* zlist;                // zones for the learned pass, run r at [zbeg_r, zend_r)
    unsigned* zcount;          // per run and list slice (list_count)
    int nzruns;
};

// Work units of a grid-stride launch.  MI355X dispatches workgroups to its 8
// XCDs round-robin by workgroup id, and each XCD has its own L2; with `xcd`
// set (and a grid that is a multiple of 8) XCD x's blocks take the x-th
// contiguous eighth of the units, so neighbouring zones -- which share points
// and write neighbouring point-corner slots -- meet in the same L2.
struct Units {
    int u, end, step;
    int band;   // the eighth of the units (x) with `xcd`, else 0
};
__device__ __forceinline__ Units block_units(int nunits, int xcd) {
    const int G = gridDim.x, b = blockIdx.x;
    Units r;
    if (xcd && G >= 8 && (G & 7) == 0) {
        const int x = b & 7, per = G >> 3;
        r.u = (int)((long long)nunits * x / 8) + (b >> 3);
        r.end = (int)((long long)nunits * (x + 1) / 8);
        r.step = per;
        r.band = x;
    } else {
        r.u = b;
        r.end = nunits;
        r.step = G;
        r.band = 0;
    }
    return r;
}

template <int NS>
__device__ __forceinline__ int nxt(int k) { return (k + 1 == NS) ? 0 : k + 1; }
template <int NS>
__device__ __forceinline__ int prv(int k) { return (k == 0) ? NS - 1 : k - 1; }

// Streaming accesses (each element touched once per kernel): nontemporal
// loads and stores, which a plain copy kernel measures 9% faster than
// ordinary ones on MI355X.
template <class T>
__device__ __forceinline__ T ld_nt(const T* p) { return __builtin_nontemporal_load(p); }
template <class T>
__device__ __forceinline__ void st_nt(T v, T* p) { __builtin_nontemporal_store(v, p); }

// A structured rect mesh, as the generator numbers it (the set-up checks
// every side against mapsp1 and pcpos before it sets a.rect_nzx): zone
// z = zy nzx + zx has sides 4z .. 4z + 3 from points p0, p0 + 1,
// p0 + 1 + npx, p0 + npx (npx = nzx + 1, p0 = zy npx + zx = z + zy), and the
// corner of side k is the rank-th corner of its point in ascending side
// order -- the number of the point's zones before z, of the four (zx-1,
// zy-1), (zx, zy-1), (zx-1, zy), (zx, zy) around it that exist. In a batch's
// arena z counts from the mesh's first side a.rect_so / 4 and p0 is offset
// by its first point a.rect_po (both 0 for a single problem).
struct RectZone {
    int p0, zx, zy;
};
__device__ __forceinline__ RectZone rect_zone(const FusedArgs& a, int z) {
    RectZone r;
    r.zy = (int)(((unsigned long long)(unsigned)z * a.rect_m) >> a.rect_sh);
    r.zx = z - r.zy * a.rect_nzx;
    r.p0 = z + r.zy + a.rect_po;
    return r;
}
__device__ __forceinline__ int rect_point(const FusedArgs& a, const RectZone& r, int k) {
    const int npx = a.rect_nzx + 1;
    return k == 0 ? r.p0 : (k == 1 ? r.p0 + 1 : (k == 2 ? r.p0 + 1 + npx : r.p0 + npx));
}
__device__ __forceinline__ int rect_rank(const FusedArgs& a, const RectZone& r, int k) {
    const int left = r.zx >= 1, below = r.zy >= 1;
    if (k == 0) return (left & below) + below + left;
    if (k == 1) return below + (below & (r.zx + 1 < a.rect_nzx));
    return k == 2 ? 0 : left;
}

// The points of a zone: mapsp1 of its sides sb..sb+NS-1.
template <int NS>
__device__ __forceinline__ void load_pp(const FusedArgs& a, int sb, int (&pp)[NS]) {
    if constexpr (NS == 4) {
        if (a.rect_nzx > 0) {
            const RectZone r = rect_zone(a, (sb - a.rect_so) >> 2);
#pragma unroll
            for (int k = 0; k < NS; ++k) pp[k] = rect_point(a, r, k);
            return;
        }
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) pp[k] = ld_nt(a.mapsp1 + sb + k);
}

// The corner-record slots of a zone's corners: pcpos of its sides (the
// slot is corner_slot(p, rank, k); a rect mesh's ranks are below 4).
template <int NS>
__device__ __forceinline__ void load_slots(const FusedArgs& a, int sb, int (&slot)[NS]) {
    if constexpr (NS == 4) {
        if (a.rect_nzx > 0) {
            const RectZone r = rect_zone(a, (sb - a.rect_so) >> 2);
#pragma unroll
            for (int k = 0; k < NS; ++k) slot[k] = rect_rank(a, r, k) * a.nump + rect_point(a, r, k);
            return;
        }
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) slot[k] = ld_nt(a.pcpos + sb + k);
}

// The learned pass's zone list of a run is kept in PN_LSLICES slices, one
// per band of block_units (all in slice 0 without the XCD mapping).  Slice s
// starts at its band's first zone, list_unit0 units into the run, so its
// entries fit in the band's own zone range; each slice has its own count on
// its own 128-byte line, so the flagging waves of the 8 XCDs add to 8
// counts instead of queueing on one.
#define PN_LSLICES 8
#define PN_LSTRIDE 32
__device__ __forceinline__ int list_unit0(int nunits, int s) {
    return (int)((long long)nunits * s / 8);
}
__device__ __forceinline__ unsigned* list_count(const FusedArgs& a, int run, int s) {
    return a.zcount + (run * PN_LSLICES + s) * PN_LSTRIDE;
}

// ---------------------------------------------------------------------------
// Side work forces (K1 writes them, K3 reads them) are stored pair-major
// within each unit of n <= PN_BLOCK zones of a run (sides u0 .. u0 + NS n - 1):
// sides 2j and 2j+1 of the unit's zone i sit together at u0 + 2 j n + 2 i
// (a lone last side of an odd NS at u0 + 2 j n + i), so that a wave's lanes
// (consecutive zones) cover one contiguous run per side pair and each lane
// moves a pair in one access.  SideW locates a zone in its unit.
struct SideW {
    int u0, i, n;
};
template <int NS>
__device__ __forceinline__ int sw_slot(const SideW& w, int k) {
    const int j = k >> 1;
    return w.u0 + 2 * j * w.n + (((k | 1) < NS) ? 2 * w.i + (k & 1) : w.i);
}
// A side pair's two values as one access (the types carry only the element
// alignment: gfx950 global memory takes dword-aligned multi-dword accesses).
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));
typedef double f64x2u __attribute__((ext_vector_type(2), aligned(8)));

__device__ __forceinline__ SideW side_w(int z, int zbeg, int zend, int sbeg, int ns) {
    const int zu = zbeg + ((z - zbeg) / PN_BLOCK) * PN_BLOCK;
    SideW w;
    w.u0 = sbeg + (zu - zbeg) * ns;
    w.i = z - zu;
    w.n = (zend - zu < PN_BLOCK) ? zend - zu : PN_BLOCK;
    return w;
}

// One corner's QCS results, as the two sides that meet at the corner use
// them: setCornerDiv, setQCnForce and setForceQCS's corner weight
// (src/QCS.cc:116-380) for corner k of a zone, whose point has velocity up0
// and half-step position xp0, neighbouring points' velocities u0p (previous)
// and u0n (next), edge centres xp1 = ex[k] and xp3 = ex[k-1], and edge
// lengths de1 = elen[k-1] (reciprocal r1) and de2 = elen[k] (reciprocal r2).
struct CrnQ {
    d2 q1, q2;      // cqe1, cqe2
    double w, c;    // cw and the clamped ccos
};
__device__ __forceinline__ CrnQ qcs_corner(d2 up0, d2 u0p, d2 u0n, d2 xp0, d2 xp1, d2 xp3, d2 zx, d2 zuc,
                                           double de1, double de2, double r1, double r2, double ztmp1,
                                           double g2, double zrp) {
    const d2 up1 = scl(0.5, add(up0, u0n));
    const d2 up2 = zuc;
    const d2 xp2 = zx;
    const d2 up3 = scl(0.5, add(u0p, up0));
    const double cvolume = 0.5 * cross(sub(xp2, xp0), sub(xp3, xp1));
    const d2 v1 = sub(xp3, xp0);
    const d2 v2 = sub(xp1, xp0);
    const double minelen = smin(de1, de2);
    const double cc = ((minelen < 1.e-12) ? 0. : 4. * dot(v1, v2) / (de1 * de2));
    const double div = (cross(sub(up2, up0), sub(xp3, xp1)) - cross(sub(up3, up1), sub(xp2, xp0))) /
                       (2.0 * cvolume);
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
    const double du = __builtin_sqrt(smax(dv1, dv2));
    const double cevol = (div < 0.0 ? evol : 0.);
    const double cdu = (div < 0.0 ? du : 0.);
    // setQCnForce (g2 = q2 * 0.25 * (qgamma + 1), the reference's left-to-right product)
    const double ztmp2 = g2 * cdu;
    const double zkur = ztmp2 + __builtin_sqrt(ztmp2 * ztmp2 + ztmp1 * ztmp1);
    const double mu = zkur * zrp * cevol;
    const double cm = ((div > 0.0) ? 0. : mu);
    CrnQ r;
    const d2 d1 = scl(cm, sub(up0, u0p));
    const d2 d2_ = scl(cm, sub(u0n, up0));
    r.q1 = mk(d1.x * r1, d1.y * r1);     // Vec2 / double: times the reciprocal
    r.q2 = mk(d2_.x * r2, d2_.y * r2);
    // setForceQCS's corner weight and clamped cosine
    const double csin2 = 1.0 - cc * cc;
    r.w = ((csin2 < 1.e-4) ? 0. : cvolume / csin2);
    r.c = ((csin2 < 1.e-4) ? 0. : cc);
    return r;
}

// K1, zone-per-thread: the whole predictor half of doCycle for zone z, whose
// sides are sb..sb+NS-1 (side work force slots sw).  Returns the zone's
// calcDtNew candidate.
//
// Every value is the reference's expression, but the corners and sides are
// visited in a streaming order that keeps few of them live at once: corner
// k+1's QCS, then side k's forces (which need corners k and k+1), then
// corner k's force tot[k] - tot[k-1]; corner 0 is kept for the last side.
// Each side's 1/elen is formed once and used by both corners at the side and
// by its viscous force (the reference divides a Vec2 by elen through the
// same reciprocal, src/Vec2.hh:145-149).
// What the FP64 predictor of a zone reads: its points' positions and
// velocities, its state from the previous cycle, its side fractions and its
// corners' slots.
template <int NS>
struct ZoneIn {
    d2 x[NS], u[NS];
    double zm, zr, ze, zvol, zwr;
    double mf[NS];
    int slot[NS];
};
template <int NS>
__device__ __forceinline__ void load_zone(const FusedArgs& a, int z, int sb, const int (&pp)[NS],
                                          ZoneIn<NS>& in) {
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        in.x[k] = mk(a.pxx[pp[k]], a.pxy[pp[k]]);
        in.u[k] = mk(a.pux[pp[k]], a.puy[pp[k]]);
    }
    in.zm = a.zm[z];
    in.zr = a.zr[z];
    in.ze = a.ze[z];
    in.zvol = a.zvol[z];
    in.zwr = a.zwrate[z];
#pragma unroll
    for (int k = 0; k < NS; ++k) in.mf[k] = a.smf[sb + k];
    load_slots<NS>(a, sb, in.slot);
}

template <int NS>
__device__ __forceinline__ double zone_pred_core(const FusedArgs& a, int z, SideW sw, const ZoneIn<NS>& in,
                                                 double dt, int& bad) {
    const double dth = 0.5 * dt;
    const double n = (double)NS;
    d2 u0[NS], xp[NS];
    double mf[NS];
    int slot[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        u0[k] = in.u[k];
        mf[k] = in.mf[k];
        slot[k] = in.slot[k];
    }
    const double zm = in.zm;
    const double zr0 = in.zr;
    const double ze0 = in.ze;
    const double zvol0 = in.zvol;
    const double zwr = in.zwr;
#pragma unroll
    for (int k = 0; k < NS; ++k) xp[k] = add(in.x[k], scl(dth, u0[k]));   // advPosHalf
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
    double sa[NS], el[NS], rel[NS];
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
    for (int k = 0; k < NS; ++k) {
        zdl = smin(zdl, fac * sa[k] / el[k]);
        rel[k] = 1. / el[k];
    }
    // calcRho (pred)
    const double zrp = zm / zvolp;
    // calcStateHalf (reads ZR, ZE, ZWRATE of the previous cycle, ZVOL0 = ZVOL)
    const double gm1 = a.gamma - 1.;
    const double ssmin2 = smax(a.ssmin * a.ssmin, 1.e-99);
    const double r = zr0;
    const double e = smax(ze0, 0.);
    const double p = gm1 * r * e;
    const double pre = gm1 * e;
    const double per = gm1 * r;
    const double csqd = smax(ssmin2, pre + per * p / (r * r));
    const double zss = __builtin_sqrt(csqd);
    const double minv = 1. / zm;
    const double dv = (zvolp - zvol0) * minv;
    const double bulk = r * csqd;
    const double denom = 1. + 0.5 * per * dv;
    const double src = zwr * dth * minv;
    const double zp = p + (per * src - r * bulk * dv) / denom;
    a.zp[z] = zp;
    // calcCrnrMass's corner mass
    const double mzz = zrp * zarea;
#pragma unroll
    for (int k = 0; k < NS; ++k) a.cmass[slot[k]] = mzz * 0.5 * (mf[k] + mf[prv<NS>(k)]);
    // calcSurfVecs, calcForcePGas, calcForceTTS per side (below)
    double sstmp = smax(zss, a.ssmin);
    sstmp = a.alfa * sstmp * sstmp;
    // setCornerDiv's zone velocity
    d2 zuc = mk(0., 0.);
#pragma unroll
    for (int k = 0; k < NS; ++k) zuc = add(zuc, dvs(u0[k], n));
    const double g2 = a.q2 * 0.25 * (a.qgamma + 1.0);
    const double ztmp1 = a.q1 * zss;
    // corners, sides and corner forces in streaming order
    const CrnQ c0 = qcs_corner(u0[0], u0[NS - 1], u0[1], xp[0], ex[0], ex[NS - 1], zx, zuc, el[NS - 1], el[0],
                               rel[NS - 1], rel[0], ztmp1, g2, zrp);
    CrnQ ck = c0;
    d2 tot0 = mk(0., 0.), totk = mk(0., 0.), sfw_prev = mk(0., 0.);
    double ztmp = 0.;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int kp = nxt<NS>(k);
        const CrnQ cn = (kp == 0) ? c0
                                  : qcs_corner(u0[kp], u0[k], u0[nxt<NS>(kp)], xp[kp], ex[kp], ex[k], zx, zuc,
                                               el[k], el[kp], rel[k], rel[kp], ztmp1, g2, zrp);
        // side k: pressure, TTS and viscous forces
        const d2 ss = rot_ccw(sub(ex[k], zx));
        const d2 sfp = scl(-zp, ss);
        const double vfacinv = zarea / sa[k];
        const double srho = zrp * mf[k] * vfacinv;
        const double sdp = sstmp * (srho - zrp);
        const d2 sft = scl(-sdp, ss);
        const d2 f = add(scl(ck.w, add(ck.q2, scl(ck.c, ck.q1))), scl(cn.w, add(cn.q1, scl(cn.c, cn.q2))));
        const d2 sfq = mk(f.x * rel[k], f.y * rel[k]);
        const d2 sfw = add(sfp, sfq);                    // calcWork's sf + sf2
        const d2 tot = add(add(sfp, sfq), sft);          // sumCrnrForce's sfp+sfq+sft
        // setVelDiff's side k, in ascending order from 0
        {
            const d2 dx = sub(xp[kp], xp[k]);
            const d2 du = sub(u0[kp], u0[k]);
            const double lenx = el[k];
            double dux = dot(du, dx);
            dux = (lenx > 0. ? fabs(dux) / lenx : 0.);
            ztmp = smax(ztmp, dux);
        }
        // corner k's force (point mapsp1[sb+k]) once sides k-1 and k are known
        if (k == 0) {
            tot0 = tot;
        } else {
            const d2 cf = sub(tot, totk);
            a.cfx[slot[k]] = cf.x;
            a.cfy[slot[k]] = cf.y;
        }
        totk = tot;
        // the work forces of a side pair together (pair-major slots)
        if (k & 1) {
            const int sl = sw_slot<NS>(sw, k - 1);
            const f64x2u vx = {sfw_prev.x, sfw.x}, vy = {sfw_prev.y, sfw.y};
            __builtin_nontemporal_store(vx, reinterpret_cast<f64x2u*>(a.sfwx + sl));
            __builtin_nontemporal_store(vy, reinterpret_cast<f64x2u*>(a.sfwy + sl));
        } else if (k + 1 == NS) {
            const int sl = sw_slot<NS>(sw, k);
            st_nt(sfw.x, a.sfwx + sl);
            st_nt(sfw.y, a.sfwy + sl);
        }
        sfw_prev = sfw;
        ck = cn;
    }
    {
        const d2 cf = sub(tot0, totk);
        a.cfx[slot[0]] = cf.x;
        a.cfy[slot[0]] = cf.y;
    }
    const double zdu = a.q1 * zss + 2. * a.q2 * ztmp;
    // calcDtNew for this zone
    const double fuzz = 1.e-99;
    const double cdt = smax(zdu, smax(zss, fuzz));
    return zdl * a.cfl / cdt;
}

template <int NS>
__device__ __forceinline__ double zone_pred(const FusedArgs& a, int z, int sb, SideW sw,
                                            const int (&pp)[NS], double dt, int& bad) {
    ZoneIn<NS> in;
    load_zone<NS>(a, z, sb, pp, in);      // every load of the zone first
    return zone_pred_core<NS>(a, z, sw, in, dt, bad);
}

template <int NS>
__device__ __forceinline__ double zone_pred(const FusedArgs& a, int z, int sb, SideW sw, double dt,
                                            int& bad) {
    int pp[NS];
    load_pp<NS>(a, sb, pp);
    return zone_pred<NS>(a, z, sb, sw, pp, dt, bad);
}

template <int NS>
__global__ void __launch_bounds__(PN_BLOCK) k_pred_zone(FusedArgs a, int zbeg, int zend, int sbeg,
                                                        int poff) {
    if (a.ctl->done) return;
    const double dt = a.ctl->dt;
    double dtmin = 1.e99;
    int bad = 0;
    const int nunits = (zend - zbeg + PN_BLOCK - 1) / PN_BLOCK;
    Units w = block_units(nunits, a.xcd);
    int zn = zbeg + w.u * PN_BLOCK + threadIdx.x;
    int pn[NS];
    if (w.u < w.end && zn < zend) load_pp<NS>(a, sbeg + (zn - zbeg) * NS, pn);
    for (; w.u < w.end; w.u += w.step) {
        // this zone's points were loaded by the previous iteration; load the next zone's
        const int z = zn;
        int pp[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) pp[k] = pn[k];
        zn = zbeg + (w.u + w.step) * PN_BLOCK + threadIdx.x;
        if (w.u + w.step < w.end && zn < zend) load_pp<NS>(a, sbeg + (zn - zbeg) * NS, pn);
        if (z >= zend) continue;
        const double zdthyd = zone_pred<NS>(a, z, sbeg + (z - zbeg) * NS,
                                            side_w(z, zbeg, zend, sbeg, NS), pp, dt, bad);
        // the running minimum of src/Hydro.cc:1615 (a NaN never replaces it)
        dtmin = (zdthyd < dtmin) ? zdthyd : dtmin;
    }
    const double m = block_min(dtmin);
    const int b = block_sum(bad);
    if (threadIdx.x == 0) {
        a.part_dt[poff + blockIdx.x] = m;
        a.part_bp[poff + blockIdx.x] = b;
    }
}

// ---------------------------------------------------------------------------
// K1, mixed precision (zone per thread).
//
// FP64: positions and their midpoints, the zone centre, side and zone areas
// and volumes (cross products of nearly equal coordinates), the zone EOS with
// its (zvolp - zvol0) difference, the TTS density difference srho - r, the
// corner-force difference tot[k] - tot[k-1], the corner mass, and the dt
// candidate.  FP32: edge lengths, the per-side pressure/TTS force products
// and the whole QCS chain (corner divergence, evolution factor, Kurapatenko
// coefficient, corner weights, edge viscous force).  Coordinate and velocity
// differences are formed in FP64 and rounded once.  Like every kernel here it
// is compiled without contraction, so each operation rounds on its own and
// the results do not depend on how the function is inlined (a batched
// problem gives the same bits as its own run).

struct f2 {
    float x, y;
};
__device__ __forceinline__ f2 fmk(float x, float y) { f2 v; v.x = x; v.y = y; return v; }
__device__ __forceinline__ f2 tof(d2 v) { return fmk((float)v.x, (float)v.y); }
__device__ __forceinline__ f2 fadd(f2 a, f2 b) { return fmk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ f2 fscl(float r, f2 v) { return fmk(v.x * r, v.y * r); }
// 1/x in FP64 from the hardware reciprocal and two Newton steps (within an
// ulp or two of the quotient; the mixed form's FP64 divisions use it)
__device__ __forceinline__ double drcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.);
    return __builtin_fma(r, e, r);
}

// The mixed form's FP32 products fold one multiply into a fused multiply-add
// (explicitly: the FP64 form is built without contraction, and explicit
// fmaf keeps every kernel that runs this code bit-identical).
__device__ __forceinline__ float fdot(f2 a, f2 b) { return __builtin_fmaf(a.x, b.x, a.y * b.y); }
__device__ __forceinline__ float fcross(f2 a, f2 b) { return __builtin_fmaf(a.x, b.y, -(a.y * b.x)); }
// a + r b
__device__ __forceinline__ f2 faxpy(float r, f2 b, f2 a) {
    return fmk(__builtin_fmaf(r, b.x, a.x), __builtin_fmaf(r, b.y, a.y));
}
// the hardware square root (1 ulp), without the correctly rounded expansion
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float flen(f2 v) { return fsqrt(fdot(v, v)); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fminr(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float fmaxr(float a, float b) { return (a < b) ? b : a; }

// The work forces of a zone's sides, {x, y} per side slot (sw_slot).
template <int NS>
__device__ __forceinline__ void store_sides(float* w, const SideW& sw, const f2 (&v)[NS]) {
#pragma unroll
    for (int k = 0; k < NS; k += 2) {
        float* q = w + 2 * sw_slot<NS>(sw, k);
        if (k + 1 < NS) {
            const f32x4u t = {v[k].x, v[k].y, v[k + 1].x, v[k + 1].y};
            __builtin_nontemporal_store(t, reinterpret_cast<f32x4u*>(q));
        } else {
            st_nt(v[k].x, q);
            st_nt(v[k].y, q + 1);
        }
    }
}
template <int NS>
__device__ __forceinline__ void load_sides(const float* w, const SideW& sw, f2 (&v)[NS]) {
#pragma unroll
    for (int k = 0; k < NS; k += 2) {
        const float* q = w + 2 * sw_slot<NS>(sw, k);
        if (k + 1 < NS) {
            const f32x4u t = __builtin_nontemporal_load(reinterpret_cast<const f32x4u*>(q));
            v[k] = fmk(t[0], t[1]);
            v[k + 1] = fmk(t[2], t[3]);
        } else {
            v[k] = fmk(ld_nt(q), ld_nt(q + 1));
        }
    }
}

// Closure kinds of the mixed K1:
//   CL_EXACT    the exact QCS coefficient;
//   CL_FLAG     the exact coefficient, and the zone is reported (ClCount.flag)
//               when one of its corners passes the network's gate;
//   CL_LEARNED  the network's coefficient on the corners that pass the gate
//               (the exact one elsewhere, and on zones outside the envelope);
//   CL_RECOMP   CL_LEARNED for a zone that a CL_FLAG pass has already handled:
//               stores only what the coefficient changes (corner and work
//               forces);
//   CL_CAPTURE  the exact coefficient, recording training tuples.
// The gate: a compressing corner (cdiv < 0) whose quadratic term matters,
// G du > PN_GATE (q1 zss + G du), i.e. f9 > PN_GATE.  Below it the exact
// coefficient equals its acoustic limit q1 zss + G du to within 0.1%; the
// round-off velocities of a quiescent region stay far below it.
enum { CL_EXACT = 0, CL_LEARNED = 1, CL_CAPTURE = 2, CL_FLAG = 3, CL_RECOMP = 4 };
#define PN_GATE 1.e-3f

// Zone counters of the learned form, summed per thread.
struct ClCount {
    int learned = 0;    // zones with gated corners evaluated by the network
    int fallback = 0;   // such zones outside the envelope (exact coefficient)
    bool flag = false;  // CL_FLAG: this zone has a gated corner
};

// The zone state a predictor starts from (the previous cycle's results).
struct ZoneState {
    double zm, zr, ze, zvol, zwrate;
};

// The FP64 half of the mixed predictor of a zone: half-step positions,
// centre, areas and volumes, edge lengths and characteristic length, the
// zone EOS and the pressure and TTS side forces, with positions and
// velocities moved into the zone's FP32 frame for the QCS half.
template <int NS>
struct MixedFrame {
    float mf[NS];     // FP32 side mass fractions
    int slot[NS];     // corner-record slots
    f2 ul[NS];        // point velocities relative to the zone-mean velocity zuc
    f2 xl[NS];        // half-step positions relative to the zone centre
    f2 exl[NS];       // edge midpoints, same frame
    float el[NS];     // edge lengths
    f2 sfp[NS];       // pressure side forces
    f2 sft[NS];       // TTS side forces
    d2 zuc;
    double zss, zrp, zarea, r;
    float zdl;
};

// The frame of zone z (sides sb..sb+NS-1) from its points' positions x0 and
// velocities u0 and its state st; writes zp to a.zp when store_zp.
template <int NS>
__device__ __forceinline__ void mixed_frame(MixedFrame<NS>& F, const FusedArgs& a, int z, int sb,
                                            const d2 (&x0)[NS], const d2 (&u0)[NS],
                                            const ZoneState& st, double dt, int& bad, bool valid,
                                            bool store_zp) {
    const double dth = 0.5 * dt;
    const double n = (double)NS;
    d2 xp[NS];
    float (&mf)[NS] = F.mf;
    const double zm = st.zm;
    const double r = st.zr;
    const double ze0 = st.ze;
    const double zvol0 = st.zvol;
    const double zwr = st.zwrate;
    F.r = r;
#pragma unroll
    for (int k = 0; k < NS; ++k) mf[k] = ld_nt(a.fsmf + sb + k);
    load_slots<NS>(a, sb, F.slot);
#pragma unroll
    for (int k = 0; k < NS; ++k)
        xp[k] = mk(__builtin_fma(dth, u0[k].x, x0[k].x), __builtin_fma(dth, u0[k].y, x0[k].y));
    // zone-mean velocity; velocities from here on live in FP32 relative to it
    d2& zuc = F.zuc;
    zuc = mk(0., 0.);
#pragma unroll
    for (int k = 0; k < NS; ++k) zuc = add(zuc, dvs(u0[k], n));
    f2 (&ul)[NS] = F.ul;
#pragma unroll
    for (int k = 0; k < NS; ++k) ul[k] = tof(sub(u0[k], zuc));
    d2 zx = mk(0., 0.);
#pragma unroll
    for (int k = 0; k < NS; ++k) zx = add(zx, dvs(xp[k], n));
    const double third = 1. / 3.;
    double sa[NS];
    float (&el)[NS] = F.el;
    double zarea = 0., zvolp = 0.;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const d2 p1 = xp[k], p2 = xp[nxt<NS>(k)];
        const d2 e = sub(p2, p1);
        sa[k] = 0.5 * cross(e, sub(zx, p1));
        const double sv = third * sa[k] * (p1.x + p2.x + zx.x);
        zarea += sa[k];
        zvolp += sv;
        if (valid && sv <= 0.) bad += 1;
        el[k] = flen(tof(e));
    }
    F.zarea = zarea;
    // positions from here on live in FP32 relative to the zone centre (which
    // is the origin of this frame); edge midpoints follow from them
    f2 (&xl)[NS] = F.xl;
    f2 (&exl)[NS] = F.exl;
#pragma unroll
    for (int k = 0; k < NS; ++k) xl[k] = tof(sub(xp[k], zx));
#pragma unroll
    for (int k = 0; k < NS; ++k) exl[k] = fscl(0.5f, fadd(xl[k], xl[nxt<NS>(k)]));
    const float fac = (NS == 3 ? 3.f : 4.f);
    float zdl = 1.e30f;
#pragma unroll
    for (int k = 0; k < NS; ++k) zdl = fminr(zdl, fac * (float)sa[k] * frcp(el[k]));
    F.zdl = zdl;
    // zone EOS in FP64
    const double zrp = zm * drcp(zvolp);
    F.zrp = zrp;
    const double gm1 = a.gamma - 1.;
    const double ssmin2 = smax(a.ssmin * a.ssmin, 1.e-99);
    const double e = smax(ze0, 0.);
    const double pz = gm1 * r * e;
    const double per = gm1 * r;
    // pre + per pz / r^2 = (gamma - 1) e + (gamma - 1)^2 e = gamma (gamma - 1) e
    const double csqd = smax(ssmin2, a.gamma * gm1 * e);
    const double zss = __builtin_sqrt(csqd);
    F.zss = zss;
    const double minv = drcp(zm);
    const double dv = (zvolp - zvol0) * minv;
    const double bulk = r * csqd;
    const double denom = 1. + 0.5 * per * dv;
    const double src = zwr * dth * minv;
    const double zp = pz + (per * src - r * bulk * dv) * drcp(denom);
    if (valid && store_zp) a.zp[z] = zp;
    // pressure and TTS side forces
    const float zpf = (float)zp;
    double sstmp = smax(zss, a.ssmin);
    sstmp = a.alfa * sstmp * sstmp;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const f2 ss = fmk(-exl[k].y, exl[k].x);                       // rotCCW(ex - zx)
        F.sfp[k] = fscl(-zpf, ss);
        // srho - r = zrp (mf zarea - sa) / sa: the difference in FP64, the
        // division in FP32 (relative to the difference itself)
        const float drho = (float)(zrp * ((double)mf[k] * zarea - sa[k])) * frcp((float)sa[k]);
        F.sft[k] = fscl(-(float)sstmp * drho, ss);
    }
}

// Corner k's QCS quantities in the frame F (setCornerDiv,
// src/QCS.cc:116-242; corner k: xp0 = x[k], xp1 = ex[k], xp2 = zx,
// xp3 = ex[k-1]; up2 = zuc is the frame's zero velocity).  evol and du are
// the values before the compression test.
struct MixedCrn {
    float area, ccos, div, evol, du, minelen;
    f2 duav;
};
template <int NS>
__device__ __forceinline__ MixedCrn mixed_corner(const MixedFrame<NS>& F, int k, f2 zucf) {
    const f2 (&ul)[NS] = F.ul;
    const f2 (&xl)[NS] = F.xl;
    const f2 (&exl)[NS] = F.exl;
    const float (&el)[NS] = F.el;
    const int km = prv<NS>(k), kp = nxt<NS>(k);
    MixedCrn c;
    const f2 u1 = fscl(0.5f, fadd(ul[k], ul[kp]));                 // up1 - zuc
    const f2 u3 = fscl(0.5f, fadd(ul[km], ul[k]));                 // up3 - zuc
    const f2 A = fmk(-xl[k].x, -xl[k].y);                           // xp2 - xp0
    const f2 B = fmk(exl[km].x - exl[k].x, exl[km].y - exl[k].y);   // xp3 - xp1
    const f2 v1 = fmk(exl[km].x - xl[k].x, exl[km].y - xl[k].y);    // xp3 - xp0
    const f2 v2 = fmk(exl[k].x - xl[k].x, exl[k].y - xl[k].y);      // xp1 - xp0
    const f2 dxx1 = fscl(0.5f, fmk(exl[k].x - xl[k].x - exl[km].x, exl[k].y - xl[k].y - exl[km].y));
    const f2 dxx2 = fscl(0.5f, fmk(exl[km].x - xl[k].x - exl[k].x, exl[km].y - xl[k].y - exl[k].y));
    const f2 du20 = fmk(-ul[k].x, -ul[k].y);                        // up2 - up0
    const f2 du31 = fmk(u3.x - u1.x, u3.y - u1.y);                  // up3 - up1
    const f2 dva = fmk(u1.x - ul[k].x - u3.x, u1.y - ul[k].y - u3.y);   // up1+up2-up0-up3
    const f2 dvb = fmk(u3.x - ul[k].x - u1.x, u3.y - ul[k].y - u1.y);   // up2+up3-up0-up1
    c.duav = faxpy(0.25f, fadd(fadd(ul[k], u1), u3), zucf);
    const float cvol = 0.5f * fcross(A, B);
    c.area = cvol;
    const float de1 = el[km], de2 = el[k];
    const float minelen = fminr(de1, de2);
    c.minelen = minelen;
    c.ccos = (minelen < 1.e-12f) ? 0.f : 4.f * fdot(v1, v2) * frcp(de1 * de2);
    c.div = (fcross(du20, B) - fcross(du31, A)) * frcp(2.f * cvol);
    const float dx1 = flen(dxx1), dx2 = flen(dxx2);
    const float t1 = fabsf(fdot(dxx1, c.duav) * dx2);
    const float t2 = fabsf(fdot(dxx2, c.duav) * dx1);
    const float num = (t1 > t2) ? dx1 : dx2;
    const float den = (t1 > t2) ? dx2 : dx1;
    float evol = fsqrt(4.f * cvol * num * frcp(den));
    c.evol = fminr(evol, 2.f * minelen);
    c.du = fsqrt(fmaxr(fdot(dva, dva), fdot(dvb, dvb)));
    return c;
}

// The network features of corner k (closure.h); the quotients are reciprocal
// products (within an ulp or two: the network reads them in BF16).
template <int NS>
__device__ __forceinline__ void corner_features(const MixedFrame<NS>& F, int k, const MixedCrn& c,
                                                float zssf, float zrpf, float q1f, float gq,
                                                float (&f)[PN_MLP_NF]) {
    const f2 (&ul)[NS] = F.ul;
    const int km = prv<NS>(k), kp = nxt<NS>(k);
    const float inv = frcp(zssf);
    f[0] = slog((ul[k].x - ul[km].x) * inv);
    f[1] = slog((ul[k].y - ul[km].y) * inv);
    f[2] = slog((ul[kp].x - ul[k].x) * inv);
    f[3] = slog((ul[kp].y - ul[k].y) * inv);
    f[4] = slog(c.div * c.evol * inv);
    f[5] = c.evol * frcp(c.minelen);
    f[6] = zssf * frcp(zssf + flen(c.duav));
    f[7] = zrpf * frcp((float)F.r);
    f[8] = slog(c.du * inv);
    const float tl = q1f * zssf, tq = gq * c.du;
    f[9] = (tl + tq > 0.f) ? tq * frcp(tl + tq) : 0.f;
}

// The mixed-precision predictor half of zone z (sides sb..sb+NS-1), from its
// points' positions x0 and velocities u0 and its state st.  With CL !=
// CL_EXACT it must be called by all 64 lanes of a wave together (the network
// runs across the wave); `valid` false marks a lane that only takes part in
// the wave-wide steps and stores nothing.  Writes zp to a.zp.
template <int NS, int CL>
__device__ __forceinline__ double zone_pred_mixed_core(const FusedArgs& a, int z, int sb, SideW sw,
                                                       const d2 (&x0)[NS], const d2 (&u0)[NS],
                                                       const ZoneState& st, double dt, int& bad,
                                                       bool valid, ClCount& cnt) {
    MixedFrame<NS> F;
    mixed_frame<NS>(F, a, z, sb, x0, u0, st, dt, bad, valid, CL != CL_RECOMP);
    const float (&mf)[NS] = F.mf;
    const int (&slot)[NS] = F.slot;
    const f2 (&ul)[NS] = F.ul;
    const f2 (&xl)[NS] = F.xl;
    const float (&el)[NS] = F.el;
    const f2 (&sfp)[NS] = F.sfp;
    const f2 (&sft)[NS] = F.sft;
    const d2 zuc = F.zuc;
    const double zss = F.zss, zrp = F.zrp, zarea = F.zarea;
    const float zdl = F.zdl;
    // QCS corner quantities (corner k: xp0 = x[k], xp1 = ex[k], xp2 = zx,
    // xp3 = ex[k-1]; up2 = zuc is the frame's zero velocity)
    const f2 zucf = tof(zuc);
    const float gq = (float)(a.q2 * 0.25 * (a.qgamma + 1.0));
    const float q1f = (float)a.q1;
    const float zssf = (float)zss;
    const float zt1 = (float)(a.q1 * zss);
    const float zrpf = (float)zrp;
    // gate threshold on du / zss (no corner passes when G = 0)
    const float gthr = (gq > 0.f) ? PN_GATE * q1f / ((1.f - PN_GATE) * gq) : 3.4e38f;
    float carea[NS], ccos[NS], cdiv[NS], cevol[NS], cdu[NS];
    u32x4 flo[NS];            // packed network inputs, slots 0-7
    unsigned fhi[NS];         // slots 8-9
    bool inenv = true;        // every gated corner inside the envelope
    bool comp = false;        // some corner passes the gate
    unsigned gated = 0;       // bit k: corner k passes the gate
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const MixedCrn c = mixed_corner<NS>(F, k, zucf);
        const float div = c.div, evol = c.evol, du = c.du;
        carea[k] = c.area;
        ccos[k] = c.ccos;
        cdiv[k] = div;
        cevol[k] = (div < 0.f) ? evol : 0.f;
        cdu[k] = (div < 0.f) ? du : 0.f;
        if constexpr (CL == CL_FLAG || CL == CL_LEARNED || CL == CL_RECOMP) {
            // G du > PN_GATE (q1 zss + G du)  <=>  du > gthr zss
            if (div < 0.f && du > gthr * zssf) {
                gated |= 1u << k;
                comp = true;
            }
        }
        if constexpr (CL == CL_LEARNED || CL == CL_RECOMP || CL == CL_CAPTURE) {
            float f[PN_MLP_NF];
            corner_features<NS>(F, k, c, zssf, zrpf, q1f, gq, f);
            if constexpr (CL == CL_LEARNED || CL == CL_RECOMP) {
                if (gated & (1u << k)) {
#pragma unroll
                    for (int j = 0; j < PN_MLP_NF; ++j)
                        inenv = inenv && (f[j] >= a.pack->lo[j]) && (f[j] <= a.pack->hi[j]);
                }
            }
            if constexpr (CL == CL_LEARNED || CL == CL_RECOMP) {
                u32x4 hi;
                pack_features(f, flo[k], hi);
                fhi[k] = hi[0];
            } else {
                // a training tuple: features and the exact coefficients
                const float tl = q1f * zssf, tq = gq * du;
                const uint64_t m = __ballot(valid && div < 0.f);
                if (m) {
                    unsigned long long base = 0;
                    if ((threadIdx.x & 63) == 0) base = atomicAdd(a.cap_count, (unsigned long long)__popcll(m));
                    base = __shfl(base, 0, 64);
                    const uint64_t below = m & ((1ull << (threadIdx.x & 63)) - 1ull);
                    const unsigned long long slot = base + (unsigned long long)__popcll(below);
                    if (valid && div < 0.f && slot < (unsigned long long)a.cap) {
                        const float s2 = fsqrt(tl * tl + tq * tq);
                        float* o = a.cap_buf + slot * PN_MLP_TUPLE;
#pragma unroll
                        for (int j = 0; j < PN_MLP_NF; ++j) o[j] = f[j];
                        o[PN_MLP_NF] = (s2 > 0.f) ? tl / s2 : 1.f;
                        o[PN_MLP_NF + 1] = (s2 > 0.f) ? 1.f + tq / s2 : 1.f;
                    }
                }
            }
        }
    }
    f2 cqe1[NS], cqe2[NS];
    float crmu[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const float zt2 = gq * cdu[k];
        const float zkur = zt2 + fsqrt(zt2 * zt2 + zt1 * zt1);
        crmu[k] = (cdiv[k] > 0.f) ? 0.f : zkur * zrpf * cevol[k];
    }
    if constexpr (CL == CL_FLAG) cnt.flag = comp;
    if constexpr (CL == CL_RECOMP) {
        // The wave's corners that need the network, packed into as few
        // evaluations as possible: corner k of a lane goes to slot off_k +
        // its rank among the lanes needing corner k, in an LDS row of the
        // wave's own; every evaluation takes 64 slots, and its results come
        // back through the same slots.  A corner's result does not depend on
        // the other corners of its evaluation (each is its own MFMA column),
        // so the bits are those of the corner-by-corner evaluation.
        __shared__ uint32_t xw_all[PN_BLOCK / 64][NS * 64][5];
        uint32_t (*xw)[5] = xw_all[threadIdx.x >> 6];
        const int lane = threadIdx.x & 63;
        const uint64_t below = (1ull << lane) - 1ull;
        const bool use = valid && comp && inenv;
        if (valid && comp) {
            if (inenv) cnt.learned += 1;
            else cnt.fallback += 1;
        }
        int dest[NS];
        int total = 0;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const bool need = use && ((gated >> k) & 1u);
            const uint64_t m = __ballot(need);
            dest[k] = need ? total + (int)__popcll(m & below) : -1;
            total += (int)__popcll(m);
            if (need) {
#pragma unroll
                for (int j = 0; j < 4; ++j) xw[dest[k]][j] = flo[k][j];
                xw[dest[k]][4] = fhi[k];
            }
        }
        // LDS accesses of one wave are performed in order; the fence keeps
        // the compiler from moving them across
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // the lane's weight fragments, read once for all of the wave's
        // evaluations
        MlpRegs wr;
        if (total > 0) mlp_load_regs(a.pack, lane, wr);
        for (int e = 0; e < total; e += 64) {
            const int sl = e + lane;
            const bool act = sl < total;
            const uint64_t am = __ballot(act);
            u32x4 lo, hi;
#pragma unroll
            for (int j = 0; j < 4; ++j) lo[j] = act ? xw[sl][j] : 0u;
            hi[0] = act ? xw[sl][4] : 0u;
            hi[1] = pack_bf16(1.f, 0.f);
            hi[2] = 0u;
            hi[3] = 0u;
            float y0, y1;
            mlp_wave_w<MlpRegs, true>(wr, lo, hi, (uint32_t)am != 0u, (am >> 32) != 0u, y0, y1);
            if (act) {
                xw[sl][0] = __float_as_uint(y0);
                xw[sl][1] = __float_as_uint(y1);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            if (dest[k] >= 0) {
                const float y0 = __uint_as_float(xw[dest[k]][0]);
                const float y1 = __uint_as_float(xw[dest[k]][1]);
                crmu[k] = (q1f * y0 * zssf + gq * y1 * cdu[k]) * zrpf * cevol[k];
            }
        }
    }
    if constexpr (CL == CL_LEARNED) {
        const bool use = valid && comp && inenv;
        if (valid && comp) {
            if (inenv) cnt.learned += 1;
            else cnt.fallback += 1;
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const bool need = use && ((gated >> k) & 1u);
            const uint64_t m = __ballot(need);
            if (m) {
                u32x4 hi;
                hi[0] = fhi[k];
                hi[1] = pack_bf16(1.f, 0.f);
                hi[2] = 0u;
                hi[3] = 0u;
                float y0, y1;
                mlp_wave_w<MlpPackSrc, true>(MlpPackSrc{a.pack, (int)(threadIdx.x & 63)}, flo[k], hi,
                                             (uint32_t)m != 0u, (m >> 32) != 0u, y0, y1);
                if (need) crmu[k] = (q1f * y0 * zssf + gq * y1 * cdu[k]) * zrpf * cevol[k];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int km = prv<NS>(k), kp = nxt<NS>(k);
        cqe1[k] = fscl(crmu[k] * frcp(el[km]), fmk(ul[k].x - ul[km].x, ul[k].y - ul[km].y));
        cqe2[k] = fscl(crmu[k] * frcp(el[k]), fmk(ul[kp].x - ul[k].x, ul[kp].y - ul[k].y));
    }
    float cw[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const float csin2 = __builtin_fmaf(-ccos[k], ccos[k], 1.f);
        cw[k] = (csin2 < 1.e-4f) ? 0.f : carea[k] * frcp(csin2);
        ccos[k] = (csin2 < 1.e-4f) ? 0.f : ccos[k];
    }
    d2 tot[NS];
    f2 sfwk[NS];
    float dux_max = 0.f;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int kp = nxt<NS>(k);
        const f2 fq = faxpy(cw[k], faxpy(ccos[k], cqe1[k], cqe2[k]),
                            fscl(cw[kp], faxpy(ccos[kp], cqe2[kp], cqe1[kp])));
        const f2 sfw = faxpy(frcp(el[k]), fq, sfp[k]);
        sfwk[k] = sfw;
        tot[k] = mk((double)sfw.x + (double)sft[k].x, (double)sfw.y + (double)sft[k].y);
        // setVelDiff
        const f2 du = fmk(ul[kp].x - ul[k].x, ul[kp].y - ul[k].y);
        const f2 dx = fmk(xl[kp].x - xl[k].x, xl[kp].y - xl[k].y);
        const float dux = fabsf(fdot(du, dx)) * frcp(el[k]);
        dux_max = fmaxr(dux_max, (el[k] > 0.f) ? dux : 0.f);
    }
    if (valid) store_sides<NS>(a.fsfw, sw, sfwk);
    const double zdu = a.q1 * zss + 2. * a.q2 * (double)dux_max;
    const double mzz = zrp * zarea;
    if (valid) {
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int km = prv<NS>(k);
            const d2 cf = sub(tot[k], tot[km]);
            a.fcrn[3 * slot[k]] = (float)cf.x;
            a.fcrn[3 * slot[k] + 1] = (float)cf.y;
            if constexpr (CL != CL_RECOMP)
                a.fcrn[3 * slot[k] + 2] = (float)(mzz * 0.5 * ((double)mf[k] + (double)mf[km]));
        }
    }
    const double cdt = smax(zdu, smax(zss, 1.e-99));
    return (double)zdl * a.cfl * drcp(cdt);
}

// The predictor of zone z whose points are pp (mapsp1 of its sides); every
// load of the zone is issued before the first use of a loaded value.  zre:
// the control block's zre_valid, read once by the caller.
template <int NS>
__device__ __forceinline__ void load_zone_mixed(const FusedArgs& a, int z, const int (&pp)[NS], bool zre,
                                                d2 (&x0)[NS], d2 (&u0)[NS], ZoneState& st) {
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        x0[k] = mk(a.pxx[pp[k]], a.pxy[pp[k]]);
        u0[k] = mk(a.pux[pp[k]], a.puy[pp[k]]);
    }
    st.zm = a.zm[z];
    st.zvol = a.zvol[z];
    st.zwrate = a.zwrate[z];
    if (zre) {
        st.zr = a.zr[z];
        st.ze = a.ze[z];
    } else {
        // the mixed corrector stored neither; these are its quotients
        st.zr = st.zm / st.zvol;
        st.ze = a.zetot[z] / (st.zm + 1.e-99);
    }
}

template <int NS, int CL>
__device__ __forceinline__ double zone_pred_mixed(const FusedArgs& a, int z, int sb, SideW sw,
                                                  const int (&pp)[NS], double dt, int& bad,
                                                  bool valid, ClCount& cnt, bool zre) {
    d2 x0[NS], u0[NS];
    ZoneState st;
    load_zone_mixed<NS>(a, z, pp, zre, x0, u0, st);
    return zone_pred_mixed_core<NS, CL>(a, z, sb, sw, x0, u0, st, dt, bad, valid, cnt);
}

template <int NS, int CL>
__device__ __forceinline__ double zone_pred_mixed(const FusedArgs& a, int z, int sb, SideW sw,
                                                  double dt, int& bad, bool valid, ClCount& cnt,
                                                  bool zre) {
    int pp[NS];
    load_pp<NS>(a, sb, pp);
    return zone_pred_mixed<NS, CL>(a, z, sb, sw, pp, dt, bad, valid, cnt, zre);
}

// K1, mixed precision.  With FLAG, zones with a gated corner are appended to
// their block's band slice of the run's zone list (list_count) for the
// learned pass; their order in the list is not fixed, their results are.
// With REDO it runs only when the pipelined cycle's finalize asked for the
// predictor to be redone (Ctl::redo).
template <int NS, bool FLAG, bool REDO = false>
__global__ void __launch_bounds__(PN_BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
k_pred_mixed(FusedArgs a, int zbeg, int zend, int sbeg, int poff, int run) {
    if (a.ctl->done) return;
    if (REDO && !a.ctl->redo) return;
    const double dt = a.ctl->dt;
    const bool zre = a.ctl->zre_valid != 0;   // read once: the loop stores
    double dtmin = 1.e99;
    int bad = 0;
    const int nunits = (zend - zbeg + PN_BLOCK - 1) / PN_BLOCK;
    Units w = block_units(nunits, a.xcd);
    int zn = zbeg + w.u * PN_BLOCK + threadIdx.x;
    int pn[NS];
    if (w.u < w.end && zn < zend) load_pp<NS>(a, sbeg + (zn - zbeg) * NS, pn);
    for (; w.u < w.end; w.u += w.step) {
        // this zone's points were loaded by the previous iteration; load the next zone's
        const int z = zn;
        int pp[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) pp[k] = pn[k];
        zn = zbeg + (w.u + w.step) * PN_BLOCK + threadIdx.x;
        if (w.u + w.step < w.end && zn < zend) load_pp<NS>(a, sbeg + (zn - zbeg) * NS, pn);
        ClCount cnt;
        if (z < zend) {
            const double zdthyd = zone_pred_mixed<NS, FLAG ? CL_FLAG : CL_EXACT>(
                a, z, sbeg + (z - zbeg) * NS, side_w(z, zbeg, zend, sbeg, NS), pp, dt, bad, true, cnt,
                zre);
            dtmin = (zdthyd < dtmin) ? zdthyd : dtmin;
        }
        if constexpr (FLAG) {
            const uint64_t m = __ballot(cnt.flag);
            if (m) {
                unsigned base = 0;
                if ((threadIdx.x & 63) == 0) base = atomicAdd(list_count(a, run, w.band), (unsigned)__popcll(m));
                base = __shfl(base, 0, 64);
                if (cnt.flag) {
                    const uint64_t below = m & ((1ull << (threadIdx.x & 63)) - 1ull);
                    a.zlist[zbeg + list_unit0(nunits, w.band) * PN_BLOCK + base + (unsigned)__popcll(below)] = z;
                }
            }
        }
    }
    const double m = block_min(dtmin);
    const int b = block_sum(bad);
    if (threadIdx.x == 0) {
        a.part_dt[poff + blockIdx.x] = m;
        a.part_bp[poff + blockIdx.x] = b;
    }
}

// Learned pass over the zones a flagging K1 listed for this run: recomputes
// each listed zone with the network's coefficient on its gated corners and
// rewrites its corner and work forces.  The packed weights are staged in LDS
// once per block.  Every lane takes part in every unit (the network runs
// across the wave); lanes past the list's end repeat its first zone and
// store nothing.
template <int NS>
__global__ void __launch_bounds__(PN_BLOCK) __attribute__((amdgpu_waves_per_eu(2)))
k_pred_list(FusedArgs a, int zbeg, int zend, int sbeg, int run) {
    // The pass is the latency of a few waves, so everything it starts from
    // (done flag, dt, the slices' counts, the weights) is loaded at once,
    // before the first wait; the weights go to LDS before the exit test.
    const int done = a.ctl->done;
    const double dt = a.ctl->dt;
    const bool zre = a.ctl->zre_valid != 0;   // read once: the loop stores
    int lcnt[PN_LSLICES];
#pragma unroll
    for (int s = 0; s < PN_LSLICES; ++s) lcnt[s] = (int)*list_count(a, run, s);
    constexpr int NQ = (int)(sizeof(ClosurePack) / 16);
    constexpr int NQT = (NQ + PN_BLOCK - 1) / PN_BLOCK;
    __shared__ ClosurePack wp;
    {
        const uint4* src = reinterpret_cast<const uint4*>(a.pack);
        uint4* dst = reinterpret_cast<uint4*>(&wp);
        // indices past the end repeat the last quad (the same value to the
        // same LDS address), so there is no branch between the loads
#pragma unroll
        for (int j = 0; j < NQT; ++j) {
            const int i = min((int)threadIdx.x + j * PN_BLOCK, NQ - 1);
            dst[i] = src[i];
        }
    }
    // the slices' units of listed zones (lu0[s] the first)
    const int runits = (zend - zbeg + PN_BLOCK - 1) / PN_BLOCK;
    int lu0[PN_LSLICES + 1];
    lu0[0] = 0;
#pragma unroll
    for (int s = 0; s < PN_LSLICES; ++s) lu0[s + 1] = lu0[s] + (lcnt[s] + PN_BLOCK - 1) / PN_BLOCK;
    const int nunits = lu0[PN_LSLICES];
    Units w = block_units(nunits, a.xcd);
    if (done || w.u >= w.end) return;     // nothing listed for this block
    __syncthreads();
    FusedArgs b = a;
    b.pack = &wp;
    int bad = 0;
    ClCount cnt;
    for (; w.u < w.end; w.u += w.step) {
        // the slice holding unit w.u (constant indices only: the arrays stay in registers)
        int s = 0, sc = lcnt[0], su0 = 0;
#pragma unroll
        for (int t = 1; t < PN_LSLICES; ++t)
            if (lu0[t] <= w.u) {
                s = t;
                sc = lcnt[t];
                su0 = lu0[t];
            }
        const int i = (w.u - su0) * PN_BLOCK + threadIdx.x;
        const bool valid = i < sc;
        const int z = a.zlist[zbeg + list_unit0(runits, s) * PN_BLOCK + (valid ? i : 0)];
        (void)zone_pred_mixed<NS, CL_RECOMP>(b, z, sbeg + (z - zbeg) * NS, side_w(z, zbeg, zend, sbeg, NS),
                                             dt, bad, valid, cnt, zre);
    }
    const int nl = block_sum(cnt.learned);
    const int nf = block_sum(cnt.fallback);
    if (threadIdx.x == 0 && (nl | nf)) {
        atomicAdd(a.cl_stats, (unsigned long long)nl);
        atomicAdd(a.cl_stats + 1, (unsigned long long)nf);
    }
}

// Mixed K1 with the learned closure or with tuple capture: every lane of a
// wave runs every unit (lanes past the run's end work on a copy of its last
// zone and store nothing), because the network and the capture slots are
// computed across the wave.
template <int NS, int CL>
__global__ void __launch_bounds__(PN_BLOCK) __attribute__((amdgpu_waves_per_eu(2)))
k_pred_cl(FusedArgs a, int zbeg, int zend, int sbeg, int poff) {
    if (a.ctl->done) return;
    const double dt = a.ctl->dt;
    const bool zre = a.ctl->zre_valid != 0;   // read once: the loop stores
    double dtmin = 1.e99;
    int bad = 0;
    ClCount cnt;
    const int nunits = (zend - zbeg + PN_BLOCK - 1) / PN_BLOCK;
    for (Units w = block_units(nunits, a.xcd); w.u < w.end; w.u += w.step) {
        const int z = zbeg + w.u * PN_BLOCK + threadIdx.x;
        const bool valid = z < zend;
        const int zc = valid ? z : zend - 1;
        const double zdthyd = zone_pred_mixed<NS, CL>(a, zc, sbeg + (zc - zbeg) * NS,
                                                      side_w(zc, zbeg, zend, sbeg, NS), dt, bad, valid,
                                                      cnt, zre);
        if (valid) dtmin = (zdthyd < dtmin) ? zdthyd : dtmin;
    }
    const double m = block_min(dtmin);
    const int b = block_sum(bad);
    if constexpr (CL == CL_LEARNED) {
        const int nl = block_sum(cnt.learned);
        const int nf = block_sum(cnt.fallback);
        if (threadIdx.x == 0 && (nl | nf)) {
            atomicAdd(a.cl_stats, (unsigned long long)nl);
            atomicAdd(a.cl_stats + 1, (unsigned long long)nf);
        }
    }
    if (threadIdx.x == 0) {
        a.part_dt[poff + blockIdx.x] = m;
        a.part_bp[poff + blockIdx.x] = b;
    }
}

// ---------------------------------------------------------------------------
// K1, corner-per-lane.

// Corner group of a lane: the NS lanes holding one zone's corners.
template <int NS>
struct Grp {
    int base;   // lane of corner 0
    int k;      // this lane's corner
    __device__ __forceinline__ double at(double v, int j) const { return __shfl(v, base + j, 64); }
    __device__ __forceinline__ double nx(double v) const {
        return __shfl(v, base + (k + 1 == NS ? 0 : k + 1), 64);
    }
    __device__ __forceinline__ double pv(double v) const {
        return __shfl(v, base + (k == 0 ? NS - 1 : k - 1), 64);
    }
    __device__ __forceinline__ d2 at(d2 v, int j) const { return mk(at(v.x, j), at(v.y, j)); }
    __device__ __forceinline__ d2 nx(d2 v) const { return mk(nx(v.x), nx(v.y)); }
    __device__ __forceinline__ d2 pv(d2 v) const { return mk(pv(v.x), pv(v.y)); }
    // sum over the group's corners in ascending order, starting from 0
    __device__ __forceinline__ double sum(double v) const {
        double acc = 0.;
#pragma unroll
        for (int j = 0; j < NS; ++j) acc += at(v, j);
        return acc;
    }
    __device__ __forceinline__ d2 sum(d2 v) const {
        d2 acc = mk(0., 0.);
#pragma unroll
        for (int j = 0; j < NS; ++j) acc = add(acc, at(v, j));
        return acc;
    }
};

// One corner of zone z (sides s - k .. s - k + NS - 1); returns the zone's
// calcDtNew candidate on the group's corner 0 and 1e99 elsewhere.
template <int NS>
__device__ __forceinline__ double corner_pred(const FusedArgs& a, const Grp<NS>& g, int z, int s,
                                              int sws, bool live, double dt, int& bad) {
    const double dth = 0.5 * dt;
    const double n = (double)NS;
    // advPosHalf at this corner's point
    const int p = a.mapsp1[s];
    const d2 x0 = mk(a.pxx[p], a.pxy[p]);
    const d2 u0 = mk(a.pux[p], a.puy[p]);
    const d2 xp = add(x0, scl(dth, u0));
    const double mf = a.smf[s];
    const d2 xpn = g.nx(xp), u0n = g.nx(u0), u0p = g.pv(u0);
    // calcCtrs
    const d2 ex = scl(0.5, add(xp, xpn));
    const d2 zx = g.sum(dvs(xp, n));
    const d2 exp_ = g.pv(ex);
    // calcVols, calcEdgeLen
    const double third = 1. / 3.;
    const double sa = 0.5 * cross(sub(xpn, xp), sub(zx, xp));
    const double sv = third * sa * (xp.x + xpn.x + zx.x);
    const double zarea = g.sum(sa);
    const double zvolp = g.sum(sv);
    if (live && sv <= 0.) bad += 1;
    const double el = len(sub(xpn, xp));
    const double elp = g.pv(el);
    // calcCharLen: min over the zone's sides in ascending order from 1e99
    const double fac = (NS == 3 ? 3. : 4.);
    const double sdl = fac * sa / el;
    double zdl = 1.e99;
#pragma unroll
    for (int j = 0; j < NS; ++j) zdl = smin(zdl, g.at(sdl, j));
    // calcRho (pred)
    const double zm = a.zm[z];
    const double zrp = zm / zvolp;
    // calcStateHalf (reads ZR, ZE, ZWRATE of the previous cycle, ZVOL0 = ZVOL)
    const double gm1 = a.gamma - 1.;
    const double ssmin2 = smax(a.ssmin * a.ssmin, 1.e-99);
    const double r = a.zr[z];
    const double e = smax(a.ze[z], 0.);
    const double pp = gm1 * r * e;
    const double pre = gm1 * e;
    const double per = gm1 * r;
    const double csqd = smax(ssmin2, pre + per * pp / (r * r));
    const double zss = __builtin_sqrt(csqd);
    const double minv = 1. / zm;
    const double dv = (zvolp - a.zvol[z]) * minv;
    const double bulk = r * csqd;
    const double denom = 1. + 0.5 * per * dv;
    const double src = a.zwrate[z] * dth * minv;
    const double zp = pp + (per * src - r * bulk * dv) / denom;
    // calcSurfVecs, calcForcePGas, calcForceTTS
    const d2 ssurf = rot_ccw(sub(ex, zx));
    const d2 sfp = scl(-zp, ssurf);
    double sstmp = smax(zss, a.ssmin);
    sstmp = a.alfa * sstmp * sstmp;
    const double vfacinv = zarea / sa;
    const double srho = zrp * mf * vfacinv;
    const double sdp = sstmp * (srho - zrp);
    const d2 sft = scl(-sdp, ssurf);
    // setCornerDiv for corner k (s = ss3[c] is side k-1, s2 = c is side k)
    const d2 zuc = g.sum(dvs(u0, n));
    const d2 up0 = u0;
    const d2 xp0 = xp;
    const d2 up1 = scl(0.5, add(up0, u0n));
    const d2 xp1 = ex;
    const d2 up2 = zuc;
    const d2 xp2 = zx;
    const d2 up3 = scl(0.5, add(u0p, up0));
    const d2 xp3 = exp_;
    const double cvolume = 0.5 * cross(sub(xp2, xp0), sub(xp3, xp1));
    const d2 v1 = sub(xp3, xp0);
    const d2 v2 = sub(xp1, xp0);
    const double de1 = elp;
    const double de2 = el;
    const double minelen = smin(de1, de2);
    double ccos = ((minelen < 1.e-12) ? 0. : 4. * dot(v1, v2) / (de1 * de2));
    const double cdiv = (cross(sub(up2, up0), sub(xp3, xp1)) - cross(sub(up3, up1), sub(xp2, xp0))) /
                        (2.0 * cvolume);
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
    evol = (cdiv < 0.0 ? evol : 0.);
    du = (cdiv < 0.0 ? du : 0.);
    // setQCnForce
    const double gammap1 = a.qgamma + 1.0;
    const double ztmp2 = a.q2 * 0.25 * gammap1 * du;
    const double ztmp1 = a.q1 * zss;
    const double zkur = ztmp2 + __builtin_sqrt(ztmp2 * ztmp2 + ztmp1 * ztmp1);
    const double mu = zkur * zrp * evol;
    const double crmu = ((cdiv > 0.0) ? 0. : mu);
    const d2 cqe1 = dvs(scl(crmu, sub(u0, u0p)), elp);
    const d2 cqe2 = dvs(scl(crmu, sub(u0n, u0)), el);
    // setForceQCS: side k joins corner k (c1) and corner k+1 (c2)
    const double csin2 = 1.0 - ccos * ccos;
    const double cw = ((csin2 < 1.e-4) ? 0. : cvolume / csin2);
    ccos = ((csin2 < 1.e-4) ? 0. : ccos);
    const double cw2 = g.nx(cw), ccos2 = g.nx(ccos);
    const d2 cqe12 = g.nx(cqe1), cqe22 = g.nx(cqe2);
    const d2 fq = add(scl(cw, add(cqe2, scl(ccos, cqe1))), scl(cw2, add(cqe12, scl(ccos2, cqe22))));
    const d2 sfq = dvs(fq, el);
    // setVelDiff: max over the zone's sides from 0
    const d2 dxs = sub(xpn, xp);
    const d2 dus = sub(u0n, u0);
    double dux = dot(dus, dxs);
    dux = (el > 0. ? fabs(dux) / el : 0.);
    double ztmp = 0.;
#pragma unroll
    for (int j = 0; j < NS; ++j) ztmp = smax(ztmp, g.at(dux, j));
    const double zdu = a.q1 * zss + 2. * a.q2 * ztmp;
    // sumCrnrForce's corner force, calcCrnrMass's corner mass, calcWork's
    // side force
    const d2 tot = add(add(sfp, sfq), sft);
    const d2 cf = sub(tot, g.pv(tot));
    const double mfp = g.pv(mf);
    const double cmass = zrp * zarea * 0.5 * (mf + mfp);
    if (live) {
        const int slot = a.pcpos[s];
        a.cfx[slot] = cf.x;
        a.cfy[slot] = cf.y;
        a.cmass[slot] = cmass;
        const d2 sfw = add(sfp, sfq);
        a.sfwx[sws] = sfw.x;
        a.sfwy[sws] = sfw.y;
        if (g.k == 0) a.zp[z] = zp;
    }
    const double fuzz = 1.e-99;
    const double cdt = smax(zdu, smax(zss, fuzz));
    const double zdthyd = zdl * a.cfl / cdt;
    return (live && g.k == 0) ? zdthyd : 1.e99;
}

template <int NS>
__global__ void __launch_bounds__(PN_BLOCK) k_pred_corner(FusedArgs a, int zbeg, int zend, int sbeg,
                                                          int poff) {
    if (a.ctl->done) return;
    constexpr int ZPW = 64 / NS;                      // zones per wave
    constexpr int ZPB = ZPW * (PN_BLOCK / 64);        // zones per block tile
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    Grp<NS> g;
    const int gi = lane / NS;
    g.k = lane - gi * NS;
    g.base = gi * NS;
    const double dt = a.ctl->dt;
    double dtmin = 1.e99;
    int bad = 0;
    const int nunits = (zend - zbeg + ZPB - 1) / ZPB;
    for (Units w = block_units(nunits, a.xcd); w.u < w.end; w.u += w.step) {
        const int zw = zbeg + w.u * ZPB + wave * ZPW + gi;
        const bool live = (gi < ZPW) && (zw < zend);
        const int z = live ? zw : zbeg;               // idle lanes read a valid zone
        const SideW sw = side_w(z, zbeg, zend, sbeg, NS);
        const double zdthyd = corner_pred<NS>(a, g, z, sbeg + (z - zbeg) * NS + g.k,
                                              sw_slot<NS>(sw, g.k), live, dt, bad);
        dtmin = (zdthyd < dtmin) ? zdthyd : dtmin;
    }
    const double m = block_min(dtmin);
    const int b = block_sum(bad);
    if (threadIdx.x == 0) {
        a.part_dt[poff + blockIdx.x] = m;
        a.part_bp[poff + blockIdx.x] = b;
    }
}

// ---------------------------------------------------------------------------
// K2

// point update shared by the regular and the heavy-point paths
// A point's state as K2 reads it.
struct PointIn {
    d2 x0, u0;
    unsigned bits;
};
__device__ __forceinline__ PointIn load_point(const FusedArgs& a, int p) {
    PointIn q;
    q.x0 = mk(a.pxx[p], a.pxy[p]);
    q.u0 = mk(a.pux[p], a.puy[p]);
    q.bits = a.pbc[p];
    return q;
}

// T = float (the mixed form): calcWork needs (pu0 + pu) only times pxp.x, so
// that product is stored in vs2 and pxpx is not written.  vs2 holds a point's
// pair {x, y} together (K3 gathers it as one 16-byte access).
template <typename T>
__device__ __forceinline__ void point_update(const FusedArgs& a, int p, const PointIn& q,
                                             double mass, d2 f, double dt) {
    const double dth = 0.5 * dt;
    const d2 x0 = q.x0;
    const d2 uin = q.u0;
    d2 u0 = uin;
    const unsigned bits = q.bits;
    if (bits) {
        for (int b = 0; b < a.nbcs; ++b) {
            if (bits & (1u << b)) {
                const d2 v = mk(a.bcv[b], a.bcv[8 + b]);
                u0 = project(u0, v);
                f = project(f, v);
            }
        }
    }
    const double fuzz = 1.e-99;
    const d2 acc = dvs(f, smax(mass, fuzz));
    const d2 u = add(u0, scl(dt, acc));
    const d2 x = add(x0, scl(dt * 0.5, add(u0, u)));
    const double xpx = x0.x + dth * uin.x;   // pxp.x for calcWork
    const d2 vs = add(u0, u);                // pu0 + pu for calcWork
    if constexpr (sizeof(T) == 4) {
        const f64x2u v = {vs.x * xpx, vs.y * xpx};
        *reinterpret_cast<f64x2u*>(a.vs2 + 2 * (int64_t)p) = v;
    } else {
        a.pxpx[p] = xpx;
        const f64x2u v = {vs.x, vs.y};
        *reinterpret_cast<f64x2u*>(a.vs2 + 2 * (int64_t)p) = v;
    }
    a.pxx[p] = x.x;
    a.pxy[p] = x.y;
    a.pux[p] = u.x;
    a.puy[p] = u.y;
}

// Corner values by corner slot: FP64 arrays for the plain form, FP32 records
// for the mixed-precision form (summed in FP64 either way).
template <typename T> struct Corners;
template <> struct Corners<double> {
    __device__ static double mass(const FusedArgs& a, int k) { return a.cmass[k]; }
    __device__ static double fx(const FusedArgs& a, int k) { return a.cfx[k]; }
    __device__ static double fy(const FusedArgs& a, int k) { return a.cfy[k]; }
};
template <> struct Corners<float> {
    __device__ static double mass(const FusedArgs& a, int k) { return (double)a.fcrn[3 * k + 2]; }
    __device__ static double fx(const FusedArgs& a, int k) { return (double)a.fcrn[3 * k]; }
    __device__ static double fy(const FusedArgs& a, int k) { return (double)a.fcrn[3 * k + 1]; }
};

// Point gathers, fixed BCs, acceleration and position update.
//
// A high-valence point (the pie origin has one corner per centre triangle)
// takes a whole block: the block stages the corner values in LDS and one
// thread adds them in ascending order.  Called by every thread of the block.
// The slot of position c (CSR index k) of point p's corner list (the host's
// corner_slot, with a.nump the mesh's or the batch arena's point count).
__device__ __forceinline__ int corner_slot(const FusedArgs& a, int p, int c, int k) {
    return (c < 4) ? c * a.nump + p : 4 * a.nump + k;
}

template <typename T>
__device__ __forceinline__ void point_heavy(const FusedArgs& a, int p, double dt) {
    using C = Corners<T>;
    __shared__ double sm[PN_BLOCK], sx[PN_BLOCK], sy[PN_BLOCK];
    const int kb = a.pcbeg[p], ke = a.pcbeg[p + 1];
    double mass = 0.;
    d2 f = mk(0., 0.);
    for (int c0 = kb; c0 < ke; c0 += PN_BLOCK) {
        const int k = c0 + threadIdx.x;
        if (k < ke) {
            const int sl = corner_slot(a, p, k - kb, k);
            sm[threadIdx.x] = C::mass(a, sl);
            sx[threadIdx.x] = C::fx(a, sl);
            sy[threadIdx.x] = C::fy(a, sl);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const int cnt = (ke - c0 < PN_BLOCK) ? ke - c0 : PN_BLOCK;
            for (int i = 0; i < cnt; ++i) {
                mass += sm[i];
                f = add(f, mk(sx[i], sy[i]));
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) point_update<T>(a, p, load_point(a, p), mass, f, dt);
}

// An ordinary point, one thread; points with more than PN_HEAVY corners are
// left to point_heavy.
//
// The point's state and its first four corners (most points have at most
// four) are loaded together, then summed in ascending order.
template <typename T>
__device__ __forceinline__ void point_light(const FusedArgs& a, int p, double dt) {
    using C = Corners<T>;
    const int kb = a.pcbeg[p], ke = a.pcbeg[p + 1];
    const int n = ke - kb;
    if (n > PN_HEAVY) return;
    const PointIn q = load_point(a, p);
    double m4[4];
    d2 f4[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        // every point has a corner: clamped to the last one, the load is in
        // range and unconditional (no branch between the loads)
        const int sl = ((j < n) ? j : n - 1) * a.nump + p;
        m4[j] = C::mass(a, sl);
        f4[j] = mk(C::fx(a, sl), C::fy(a, sl));
    }
    double mass = 0.;
    d2 f = mk(0., 0.);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (kb + j < ke) {
            mass += m4[j];
            f = add(f, f4[j]);
        }
    }
    for (int k = kb + 4; k < ke; ++k) {
        const int sl = 4 * a.nump + k;
        mass += C::mass(a, sl);
        f = add(f, mk(C::fx(a, sl), C::fy(a, sl)));
    }
    point_update<T>(a, p, q, mass, f, dt);
}

// K2: blocks [0, nheavy) take the heavy points, the rest one point per thread.
// Block 0 also clears the learned pass's zone counts: that pass has run
// before K2, and the next predictor appends to them.
template <bool REDO>
__device__ __forceinline__ void fold_lookahead(const FusedArgs& a, int npred);

template <typename T>
__global__ void __launch_bounds__(PN_BLOCK) k_point(FusedArgs a) {
    // The pipelined cycle's K2 has one block more: it folds the partials of
    // the previous cycle's redone predictor, if there was one, into the
    // look-ahead values -- the work of a separate one-block launch after the
    // redo, here before the zone pass overwrites the partials.
    if (a.la_npred > 0 && (int)blockIdx.x == (int)gridDim.x - 1) {
        fold_lookahead<true>(a, a.la_npred);
        return;
    }
    if (a.zcount && blockIdx.x == 0)
        for (int i = threadIdx.x; i < a.nzruns * PN_LSLICES; i += PN_BLOCK) a.zcount[i * PN_LSTRIDE] = 0u;
    if (a.ctl->done) return;
    const double dt = a.ctl->dt;
    if ((int)blockIdx.x < a.nheavy) {
        point_heavy<T>(a, a.heavy[blockIdx.x], dt);
        return;
    }
    const int p = (blockIdx.x - a.nheavy) * PN_BLOCK + threadIdx.x;
    if (p < a.npown) point_light<T>(a, p, dt);
}

// ---------------------------------------------------------------------------
// K3: corrector geometry, work, work rate, energy, density, dvol of zone z
// (sides sb..sb+NS-1, points pp); folds the bad-side count and the dvol
// ratio.  The zone's own loads are issued before the point gathers, which
// wait for the point indices.  Returns the points' positions x and the zone's
// new state st (the next cycle's predictor starts from them).  With PIPE the
// cycle's predictor pressure comes from a.zpn and is stored to a.zp.
template <int NS, typename T, bool PIPE>
__device__ __forceinline__ void corr_zone_core(const FusedArgs& a, int z, int sb, SideW sw,
                                               const int (&pp)[NS], double dt, int& bad,
                                               double& dvov, d2 (&x)[NS], ZoneState& st) {
    double et = a.zetot[z];
    const double zvol0 = a.zvol[z];
    const double zm = a.zm[z];
    const double zp = PIPE ? a.zpn[z] : a.zp[z];
    d2 sfw[NS];
    if constexpr (sizeof(T) == 4) {
        f2 v[NS];
        load_sides<NS>(a.fsfw, sw, v);
#pragma unroll
        for (int k = 0; k < NS; ++k) sfw[k] = mk((double)v[k].x, (double)v[k].y);
    } else {
#pragma unroll
        for (int k = 0; k < NS; k += 2) {
            const int sl = sw_slot<NS>(sw, k);
            if (k + 1 < NS) {
                const f64x2u vx = __builtin_nontemporal_load(reinterpret_cast<const f64x2u*>(a.sfwx + sl));
                const f64x2u vy = __builtin_nontemporal_load(reinterpret_cast<const f64x2u*>(a.sfwy + sl));
                sfw[k] = mk(vx[0], vy[0]);
                sfw[k + 1] = mk(vx[1], vy[1]);
            } else {
                sfw[k] = mk(ld_nt(a.sfwx + sl), ld_nt(a.sfwy + sl));
            }
        }
    }
    d2 vs[NS];                 // mixed form: (pu0 + pu) pxp.x
    double xp[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        x[k] = mk(a.pxx[pp[k]], a.pxy[pp[k]]);
        const f64x2u v = *reinterpret_cast<const f64x2u*>(a.vs2 + 2 * (int64_t)pp[k]);
        vs[k] = mk(v[0], v[1]);
        if constexpr (sizeof(T) == 8) xp[k] = a.pxpx[pp[k]];
    }
    const double dth = 0.5 * dt;
    const double n = (double)NS;
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
    double w = 0.;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int k2 = nxt<NS>(k);
        const d2 sftot = sfw[k];
        const double sd1 = dot(sftot, vs[k]);
        const double sd2 = dot(neg(sftot), vs[k2]);
        const double dwork = (sizeof(T) == 8) ? -dth * (sd1 * xp[k] + sd2 * xp[k2]) : -dth * (sd1 + sd2);
        et += dwork;
        w += dwork;
    }
    // calcWorkRate, calcEnergy, calcRho (corr), calcDvol
    const double dtinv = 1. / dt;
    const double dvol = zvol - zvol0;
    const double zwr = (w + zp * dvol) * dtinv;
    a.zwrate[z] = zwr;
    a.zetot[z] = et;
    const double fuzz = 1.e-99;
    const double ze = et / (zm + fuzz);
    const double zr = zm / zvol;
    if constexpr (sizeof(T) == 8) {
        // the mixed form's predictors derive zr and ze from zm, zvol, zetot
        a.ze[z] = ze;
        a.zr[z] = zr;
    }
    a.zvol[z] = zvol;
    if (PIPE) a.zp[z] = zp;
    const double zdvov = fabs((zvol - zvol0) / zvol0);
    dvov = (zdvov > dvov) ? zdvov : dvov;
    st.zm = zm;
    st.zr = zr;
    st.ze = ze;
    st.zvol = zvol;
    st.zwrate = zwr;
}

template <int NS, typename T>
__device__ __forceinline__ void corr_zone_pp(const FusedArgs& a, int z, int sb, SideW sw,
                                             const int (&pp)[NS], double dt, int& bad, double& dvov) {
    d2 x[NS];
    ZoneState st;
    corr_zone_core<NS, T, false>(a, z, sb, sw, pp, dt, bad, dvov, x, st);
}

template <int NS, typename T>
__device__ __forceinline__ void corr_zone(const FusedArgs& a, int z, int sb, SideW sw, double dt,
                                          int& bad, double& dvov) {
    int pp[NS];
    load_pp<NS>(a, sb, pp);
    corr_zone_pp<NS, T>(a, z, sb, sw, pp, dt, bad, dvov);
}

// The point indices of a thread's next zone are loaded while the current
// zone is processed.
template <int NS, typename T>
__global__ void __launch_bounds__(PN_BLOCK) k_corr(FusedArgs a, int zbeg, int zend, int sbeg,
                                                   int poff) {
    if (a.ctl->done) return;
    const double dt = a.ctl->dt;
    double dvov = 1.e-99;
    int bad = 0;
    const int nunits = (zend - zbeg + PN_BLOCK - 1) / PN_BLOCK;
    Units wu = block_units(nunits, a.xcd);
    int z = zbeg + wu.u * PN_BLOCK + threadIdx.x;
    int pp[NS];
    if (wu.u < wu.end && z < zend) load_pp<NS>(a, sbeg + (z - zbeg) * NS, pp);
    for (; wu.u < wu.end; wu.u += wu.step) {
        const int zc = z;
        int pc[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) pc[k] = pp[k];
        z = zbeg + (wu.u + wu.step) * PN_BLOCK + threadIdx.x;
        if (wu.u + wu.step < wu.end && z < zend) load_pp<NS>(a, sbeg + (z - zbeg) * NS, pp);
        if (zc < zend)
            corr_zone_pp<NS, T>(a, zc, sbeg + (zc - zbeg) * NS, side_w(zc, zbeg, zend, sbeg, NS), pc, dt,
                                bad, dvov);
    }
    const double m = block_max(dvov);
    const int b = block_sum(bad);
    if (threadIdx.x == 0) {
        a.part_dv[poff + blockIdx.x] = m;
        a.part_bc[poff + blockIdx.x] = b;
    }
}

// ---------------------------------------------------------------------------
// Pipelined mixed cycle.  The predictor of cycle n+1 needs dt(n+1), which the
// finalize of cycle n computes from this cycle's predictor dt limit (known
// before K2) and its corrector's volume-change limit (known after K3).  The
// zone pass of cycle n runs each zone's corrector and then, from the state it
// has just computed, the zone's predictor of cycle n+1 with the dt that the
// finalize gives when the volume-change limit does not bind.  The finalize
// checks that dt bit for bit; when it differs, the predictor is redone with
// the actual dt.  Either way cycle n+1 starts from the same bits as the
// unpipelined cycle.

// The control-block values the end of a cycle and the look-ahead dt start from.
struct CtlSnap {
    double time, dt, la_dtnew;
    int cycle, la_bp, done;
};
// Relaxed atomic loads: vector loads (never scalar-cache loads), issued
// back to back with no wait between them.
__device__ __forceinline__ double ld_vec(const double* p) {
    return __builtin_bit_cast(double, __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p),
                                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ int ld_vec(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Read through vector loads: the caller stores to the block later.
__device__ __forceinline__ CtlSnap snap_ctl(const Ctl* c) {
    CtlSnap v;
    v.time = ld_vec(&c->time);
    v.dt = ld_vec(&c->dt);
    v.la_dtnew = ld_vec(&c->la_dtnew);
    v.cycle = ld_vec(&c->cycle);
    v.la_bp = ld_vec(&c->la_bp);
    v.done = ld_vec(&c->done);
    return v;
}

struct DtParams {
    double cflv, dtmax, dtfac, tstop;
    int cstop;
};

// dt of the cycle after the current one as the finalize computes it
// (src/Hydro.cc:1647-1667, src/Driver.cc:213-245) when the volume-change
// limit does not bind; `last` when the current cycle ends the run.
__device__ __forceinline__ double lookahead_dt(const CtlSnap& c, const DtParams& q, bool& last) {
    const double t1 = c.time + c.dt;
    last = (c.cycle + 1 >= q.cstop || t1 >= q.tstop);
    double r = 1.e99;
    if (c.la_dtnew < r) r = c.la_dtnew;
    double dt = q.dtmax;
    const double dtrecover = q.dtfac * c.dt;
    if (dtrecover < dt) dt = dtrecover;
    if ((q.tstop - t1) < dt) dt = q.tstop - t1;
    if (r < dt) dt = r;
    return dt;
}

// Zone pass of the pipelined mixed cycle: the corrector of every zone, then
// (unless the cycle ends the run) its predictor of the next cycle, which
// writes zp to a.zpn and, with FLAG, lists zones for the learned pass.
// Partials: corrector at part_dv/part_bc, predictor at part_dt/part_bp.
template <int NS, bool FLAG>
__global__ void __launch_bounds__(PN_BLOCK) __attribute__((amdgpu_waves_per_eu(3)))
k_zone_pipe(FusedArgs a, int zbeg, int zend, int sbeg, int poff, int run, DtParams q) {
    const Ctl* c = a.ctl;
    if (c->done) return;
    const double dt = c->dt;
    CtlSnap v;                 // this kernel does not write the control block
    v.time = c->time;
    v.dt = dt;
    v.la_dtnew = c->la_dtnew;
    v.cycle = c->cycle;
    v.la_bp = 0;
    v.done = 0;
    bool last;
    const double dtn = lookahead_dt(v, q, last);
    FusedArgs b = a;
    b.zp = a.zpn;
    double dvov = 1.e-99, dtmin = 1.e99;
    int badc = 0, badp = 0;
    const int nunits = (zend - zbeg + PN_BLOCK - 1) / PN_BLOCK;
    Units w = block_units(nunits, a.xcd);
    int zn = zbeg + w.u * PN_BLOCK + threadIdx.x;
    int pn[NS];
    if (w.u < w.end && zn < zend) load_pp<NS>(a, sbeg + (zn - zbeg) * NS, pn);
    for (; w.u < w.end; w.u += w.step) {
        const int z = zn;
        int pp[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) pp[k] = pn[k];
        zn = zbeg + (w.u + w.step) * PN_BLOCK + threadIdx.x;
        if (w.u + w.step < w.end && zn < zend) load_pp<NS>(a, sbeg + (zn - zbeg) * NS, pn);
        ClCount cnt;
        if (z < zend) {
            const int sb = sbeg + (z - zbeg) * NS;
            const SideW sw = side_w(z, zbeg, zend, sbeg, NS);
            d2 u[NS];
            if (!last) {
#pragma unroll
                for (int k = 0; k < NS; ++k) u[k] = mk(a.pux[pp[k]], a.puy[pp[k]]);
            }
            d2 x[NS];
            ZoneState st;
            corr_zone_core<NS, float, true>(a, z, sb, sw, pp, dt, badc, dvov, x, st);
            if (!last) {
                const double zdthyd = zone_pred_mixed_core<NS, FLAG ? CL_FLAG : CL_EXACT>(
                    b, z, sb, sw, x, u, st, dtn, badp, true, cnt);
                dtmin = (zdthyd < dtmin) ? zdthyd : dtmin;
            }
        }
        if constexpr (FLAG) {
            const uint64_t m = __ballot(cnt.flag);
            if (m) {
                unsigned base = 0;
                if ((threadIdx.x & 63) == 0) base = atomicAdd(list_count(a, run, w.band), (unsigned)__popcll(m));
                base = __shfl(base, 0, 64);
                if (cnt.flag) {
                    const uint64_t below = m & ((1ull << (threadIdx.x & 63)) - 1ull);
                    a.zlist[zbeg + list_unit0(nunits, w.band) * PN_BLOCK + base + (unsigned)__popcll(below)] = z;
                }
            }
        }
    }
    const double mv = block_max(dvov);
    const int bc = block_sum(badc);
    const double mt = block_min(dtmin);
    const int bp = block_sum(badp);
    if (threadIdx.x == 0) {
        a.part_dv[poff + blockIdx.x] = mv;
        a.part_bc[poff + blockIdx.x] = bc;
        a.part_dt[poff + blockIdx.x] = mt;
        a.part_bp[poff + blockIdx.x] = bp;
    }
}

// ---------------------------------------------------------------------------
// K4: reductions over the block partials (min and max are exact in any
// order; NaN partials cannot occur, the per-zone folds skip NaN) and the end
// of the cycle: calcDt (src/Hydro.cc:1647-1667), checkBadSides
// (src/Mesh.cc:1531-1550), the time/cycle update (src/Driver.cc:266-295), the
// stop test and calcGlobalDt (src/Driver.cc:213-245) for the next cycle.
__device__ __forceinline__ double blk_min_fin(double v) {
    __shared__ double part[PN_FIN_BLOCK / 64];
    __syncthreads();   // a previous call's partials may still be read
    v = wave_min(v);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 1; i < PN_FIN_BLOCK / 64; ++i) v = (part[i] < v) ? part[i] : v;
    return v;
}
__device__ __forceinline__ int blk_sum_fin(int v) {
    __shared__ int part[PN_FIN_BLOCK / 64];
    __syncthreads();   // a previous call's partials may still be read
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 1; i < PN_FIN_BLOCK / 64; ++i) v += part[i];
    return v;
}

// The four folds of a finalize in one pass over the block: wave folds, one
// LDS row per wave, and thread 0 folds the rows in wave order (the result is
// thread 0's).
__device__ __forceinline__ void blk_fold_fin(double& mn, double& mx, int& bp, int& bc) {
    __shared__ double pmn[PN_FIN_BLOCK / 64], pmx[PN_FIN_BLOCK / 64];
    __shared__ int pbp[PN_FIN_BLOCK / 64], pbc[PN_FIN_BLOCK / 64];
    mn = wave_min(mn);
    mx = wave_max(mx);
    bp = wave_sum(bp);
    bc = wave_sum(bc);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        pmn[w] = mn;
        pmx[w] = mx;
        pbp[w] = bp;
        pbc[w] = bc;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 1; i < PN_FIN_BLOCK / 64; ++i) {
            mn = (pmn[i] < mn) ? pmn[i] : mn;
            mx = (pmx[i] > mx) ? pmx[i] : mx;
            bp += pbp[i];
            bc += pbc[i];
        }
}

// The end of the cycle for one problem, by one block of PN_FIN_BLOCK threads,
// from npred predictor and ncorr corrector partials.
__device__ __forceinline__ int end_cycle(Ctl* c, const CtlSnap& v, double mn, double mx, int bp,
                                         int bc, double cflv, int cstop, double dtmax,
                                         double dtfac, double tstop, int zre,
                                         double* next_dt = nullptr);

// Each thread's share of the partials (thread t folds t, t + PN_FIN_BLOCK,
// ...), loaded PN_FIN_PER per thread at a time with no branch between the
// loads.  Indices past the end read the last partial again: a repeat leaves a
// min or max unchanged and adds 0 to a count.  npred, ncorr >= 1 (every
// launch has at least one block, every problem at least one unit).
#define PN_FIN_PER 8
__device__ __forceinline__ void fold_partials(const double* part_dt, const int* part_bp, int npred,
                                              const double* part_dv, const int* part_bc, int ncorr,
                                              double& mn, int& bp, double& mx, int& bc) {
    const int n = npred > ncorr ? npred : ncorr;
    for (int base = 0; base < n; base += PN_FIN_PER * PN_FIN_BLOCK) {
        double dv[PN_FIN_PER], dd[PN_FIN_PER];
        int ip[PN_FIN_PER], ic[PN_FIN_PER];
#pragma unroll
        for (int j = 0; j < PN_FIN_PER; ++j) {
            const int i = base + (int)threadIdx.x + j * PN_FIN_BLOCK;
            const int ipd = min(i, npred - 1), icr = min(i, ncorr - 1);
            dd[j] = part_dt[ipd];
            ip[j] = part_bp[ipd];
            dv[j] = part_dv[icr];
            ic[j] = part_bc[icr];
        }
#pragma unroll
        for (int j = 0; j < PN_FIN_PER; ++j) {
            const int i = base + (int)threadIdx.x + j * PN_FIN_BLOCK;
            mn = (dd[j] < mn) ? dd[j] : mn;
            bp += (i < npred) ? ip[j] : 0;
            mx = (dv[j] > mx) ? dv[j] : mx;
            bc += (i < ncorr) ? ic[j] : 0;
        }
    }
}

__device__ __forceinline__ void finalize_block(Ctl* c, const double* part_dt, const int* part_bp,
                                               int npred, const double* part_dv, const int* part_bc,
                                               int ncorr, double cflv, int cstop, double dtmax,
                                               double dtfac, double tstop, int zre) {
    // The control block and the partials are read at once, so the block
    // waits for memory once (the partials of a run that has stopped are read
    // but not used).
    const CtlSnap v = snap_ctl(c);
    double mn = 1.e99, mx = 1.e-99;
    int bp = 0, bc = 0;
    fold_partials(part_dt, part_bp, npred, part_dv, part_bc, ncorr, mn, bp, mx, bc);
    blk_fold_fin(mn, mx, bp, bc);
    if (threadIdx.x != 0 || v.done) return;
    end_cycle(c, v, mn, mx, bp, bc, cflv, cstop, dtmax, dtfac, tstop, zre);
}

// The end of a cycle from its predictor's dt limit and bad-side count (mn,
// bp) and its corrector's dvol ratio and bad-side count (mx, bc).  v is the
// control block as read through vector loads before any store to it (a
// scalar-cache load is not ordered with a later vector store to the same
// address).
// Returns the new done flag; *next_dt (when given) receives the next dt.
// zre: whether this cycle's corrector stored zr and ze (Ctl::zre_valid).
__device__ __forceinline__ int end_cycle(Ctl* c, const CtlSnap& v, double mn, double mx, int bp,
                                         int bc, double cflv, int cstop, double dtmax,
                                         double dtfac, double tstop, int zre, double* next_dt) {
    c->zre_valid = zre;
    double r = 1.e99;
    if (mn < r) r = mn;
    const double d2v = v.dt * cflv / mx;
    if (d2v < r) r = d2v;
    c->dtnew = mn;
    c->dvovmax = mx;
    c->numsbad_pred = bp;
    c->numsbad_corr = bc;
    c->dtrec = r;
    if (bp > 0 || bc > 0) {
        c->err_cycle = v.cycle;
        c->err_pred = bp > 0 ? 1 : 0;
        c->done = 1;
        return 1;
    }
    const double time = v.time + v.dt;
    const int cycle = v.cycle + 1;
    const double dtlast = v.dt;
    double dt = dtmax;
    const double dtrecover = dtfac * dtlast;
    if (dtrecover < dt) dt = dtrecover;
    if ((tstop - time) < dt) dt = tstop - time;
    if (r < dt) dt = r;
    const int done = (cycle >= cstop || time >= tstop) ? 1 : 0;
    c->time = time;
    c->cycle = cycle;
    c->dtlast = dtlast;
    c->done = done;
    c->dt = dt;
    if (next_dt) *next_dt = dt;
    return done;
}

__global__ void __launch_bounds__(PN_FIN_BLOCK) k_finalize(FusedArgs a, int npred, int ncorr,
                                                           double cflv, int cstop, double dtmax,
                                                           double dtfac, double tstop, int zre) {
    // the learned pass of this cycle has consumed its zone lists
    if (a.zcount)
        for (int i = threadIdx.x; i < a.nzruns * PN_LSLICES; i += PN_FIN_BLOCK) a.zcount[i * PN_LSTRIDE] = 0u;
    finalize_block(a.ctl, a.part_dt, a.part_bp, npred, a.part_dv, a.part_bc, ncorr, cflv, cstop,
                   dtmax, dtfac, tstop, zre);
}

// K4 of the pipelined mixed cycle: ends cycle n from its predictor's results
// (carried in Ctl::la_*) and its corrector's partials, then accepts the
// look-ahead predictor of cycle n+1 when the dt it assumed is the dt just
// computed, or asks for it to be redone (clearing the learned pass's zone
// counts, which the redo fills again).  The learned pass's zone counts of
// cycle n move from the pending to the reported counters.
__global__ void __launch_bounds__(PN_FIN_BLOCK) k_finalize_pipe(FusedArgs a, int npred, int ncorr,
                                                                DtParams q) {
    Ctl* c = a.ctl;
    if (c->done) return;
    double mn = 1.e99, mx = 1.e-99;
    int bp = 0, bc = 0;
    fold_partials(a.part_dt, a.part_bp, npred, a.part_dv, a.part_bc, ncorr, mn, bp, mx, bc);
    blk_fold_fin(mn, mx, bp, bc);
    if (threadIdx.x != 0) return;
    if (a.cl_stats) {
        const unsigned long long nl = atomicExch(a.cl_stats + 2, 0ull);
        const unsigned long long nf = atomicExch(a.cl_stats + 3, 0ull);
        atomicAdd(a.cl_stats, nl);
        atomicAdd(a.cl_stats + 1, nf);
    }
    const CtlSnap v = snap_ctl(c);
    bool last;
    const double dtn = lookahead_dt(v, q, last);
    double dt1 = 0.;
    if (end_cycle(c, v, v.la_dtnew, mx, v.la_bp, bc, q.cflv, q.cstop, q.dtmax, q.dtfac, q.tstop, 0,
                  &dt1))
        return;
    if (!last && dt1 == dtn) {
        c->la_dtnew = mn;
        c->la_bp = bp;
        c->redo = 0;
    } else {
        c->redo = 1;
        c->la_redos += 1;
        if (a.zcount)
            for (int r = 0; r < a.nzruns * PN_LSLICES; ++r) a.zcount[r * PN_LSTRIDE] = 0u;
    }
}

// After a predictor of the pipelined cycle (the prologue's, or with REDO a
// redone one): its dt limit and bad-side count become the Ctl::la_* values
// the next finalize ends the cycle with.  One block of PN_FIN_BLOCK threads.
template <bool REDO>
__device__ __forceinline__ void fold_lookahead(const FusedArgs& a, int npred) {
    Ctl* c = a.ctl;
    if (c->done) return;
    if (REDO && !c->redo) return;
    double mn = 1.e99;
    int bp = 0;
    for (int i = threadIdx.x; i < npred; i += PN_FIN_BLOCK) {
        const double v = a.part_dt[i];
        mn = (v < mn) ? v : mn;
        bp += a.part_bp[i];
    }
    mn = blk_min_fin(mn);
    bp = blk_sum_fin(bp);
    if (threadIdx.x != 0) return;
    c->la_dtnew = mn;
    c->la_bp = bp;
    c->redo = 0;
}
template <bool REDO>
__global__ void __launch_bounds__(PN_FIN_BLOCK) k_finalize_la(FusedArgs a, int npred) {
    fold_lookahead<REDO>(a, npred);
}

FusedArgs make_args(pennant_problem* pb) {
    FusedArgs a;
    const Params& q = pb->prm;
    a.mapsp1 = pb->dm.mapsp1; a.pcbeg = pb->dm.pcbeg; a.pcpos = pb->pcpos;
    a.smf = pb->f("smf"); a.nump = pb->dm.nump; a.npown = own_points(pb); a.la_npred = 0;
    a.pxx = pb->f("px", 0); a.pxy = pb->f("px", 1); a.pux = pb->f("pu", 0); a.puy = pb->f("pu", 1);
    a.zr = pb->f("zr"); a.ze = pb->f("ze"); a.zetot = pb->f("zetot"); a.zwrate = pb->f("zwrate");
    a.zvol = pb->f("zvol"); a.zp = pb->f("zp"); a.zm = pb->f("zm");
    a.cfx = pb->cfx; a.cfy = pb->cfy; a.cmass = pb->cmass; a.sfwx = pb->sfwx; a.sfwy = pb->sfwy;
    a.fcrn = pb->fcrn; a.fsfw = pb->fsfw; a.fsmf = pb->fsmf; a.zpn = pb->zpn;
    a.vs2 = pb->vs2; a.pxpx = pb->pxpx; a.pbc = pb->pbc;
    a.heavy = pb->heavy; a.nheavy = pb->nheavy;
    a.part_dt = pb->part_d; a.part_dv = pb->part_d + pb->npart_pred;
    a.part_bp = pb->part_i; a.part_bc = pb->part_i + pb->npart_pred;
    a.ctl = pb->ctl;
    a.gamma = q.gamma; a.ssmin = q.ssmin; a.alfa = q.alfa; a.qgamma = q.qgamma;
    a.q1 = q.q1; a.q2 = q.q2; a.cfl = q.cfl;
    a.nbcs = pb->nbcs;
    a.bcv = pb->bcv_dev;
    a.xcd = pb->tune.xcd;
    a.pack = (const ClosurePack*)pb->closure;
    a.cap_buf = pb->cap_buf;
    a.cap = pb->cap;
    a.cap_count = pb->cl_counters;
    a.cl_stats = pb->cl_counters + 1;
    a.rect_nzx = (pb->tune.rect_maps && pb->rect_nzx > 0) ? pb->rect_nzx : 0;
    a.rect_m = pb->rect_m;
    a.rect_sh = pb->rect_sh;
    a.rect_so = 0;
    a.rect_po = 0;
    a.zlist = pb->zlist;
    a.zcount = pb->zcount;
    a.nzruns = (int)pb->zruns.size() / 3;
    return a;
}

__global__ void __launch_bounds__(PN_BLOCK) k_to_float(int n, const double* src, float* dst) {
    const int i = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (i < n) dst[i] = (float)src[i];
}

// zr and ze as the FP64 corrector stores them (zr = zm / zvol, ze = zetot /
// (zm + 1e-99)), where mixed cycles left them to be derived; then the flag.
__global__ void __launch_bounds__(PN_BLOCK) k_zre(FusedArgs a, int nz) {
    if (a.ctl->zre_valid) return;
    const int z = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (z < nz) {
        const double zm = a.zm[z];
        a.zr[z] = zm / a.zvol[z];
        a.ze[z] = a.zetot[z] / (zm + 1.e-99);
    }
}
__global__ void __launch_bounds__(64) k_zre_done(Ctl* c, int n) {
    for (int i = threadIdx.x; i < n; i += 64) c[i].zre_valid = 1;
}

// Small fills (kernels rather than memsets, so captured graphs hold kernel
// nodes only); one block of 64 threads.
__global__ void __launch_bounds__(64) k_zero_u32(unsigned* p, int n) {
    for (int i = threadIdx.x; i < n; i += 64) p[i] = 0u;
}
__global__ void __launch_bounds__(64) k_zero_u64(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

int grid_for(int needed, int cap) { return needed < 1 ? 1 : (needed < cap ? needed : cap); }

// blocks of each K1 / K3 launch for a run of nz zones with NS sides
int pred_grid(const pennant_problem* pb, int nz, int ns, int form) {
    const int per = (form == FORM_FP64 && pb->tune.k1_layout == 1) ? (PN_BLOCK / 64) * (64 / ns) : PN_BLOCK;
    return grid_for((nz + per - 1) / per, pb->tune.max_blocks);
}
int corr_grid(const pennant_problem* pb, int nz) {
    return grid_for(nblocks(nz), pb->tune.max_blocks);
}

template <int NS>
void launch_pred(const pennant_problem* pb, const FusedArgs& a, int zb, int ze, int sb, int poff,
                 int form, int run, hipStream_t s) {
    const dim3 grid(pred_grid(pb, ze - zb, NS, form));
    if (form == FORM_LEARNED)
        hipLaunchKernelGGL((k_pred_mixed<NS, true>), grid, dim3(PN_BLOCK), 0, s, a, zb, ze, sb, poff, run);
    else if (form == FORM_MIXED && a.cap_buf)
        hipLaunchKernelGGL((k_pred_cl<NS, CL_CAPTURE>), grid, dim3(PN_BLOCK), 0, s, a, zb, ze, sb, poff);
    else if (form == FORM_MIXED)
        hipLaunchKernelGGL((k_pred_mixed<NS, false>), grid, dim3(PN_BLOCK), 0, s, a, zb, ze, sb, poff, run);
    else if (pb->tune.k1_layout == 1)
        hipLaunchKernelGGL(k_pred_corner<NS>, grid, dim3(PN_BLOCK), 0, s, a, zb, ze, sb, poff);
    else
        hipLaunchKernelGGL(k_pred_zone<NS>, grid, dim3(PN_BLOCK), 0, s, a, zb, ze, sb, poff);
}
template <int NS>
void launch_corr(const pennant_problem* pb, const FusedArgs& a, int zb, int ze, int sb, int poff,
                 int form, hipStream_t s) {
    const dim3 grid(corr_grid(pb, ze - zb));
    if (form != FORM_FP64)
        hipLaunchKernelGGL((k_corr<NS, float>), grid, dim3(PN_BLOCK), 0, s, a, zb, ze, sb, poff);
    else
        hipLaunchKernelGGL((k_corr<NS, double>), grid, dim3(PN_BLOCK), 0, s, a, zb, ze, sb, poff);
}

// ---------------------------------------------------------------------------
// Batched problems: one launch per phase over the units of every problem.
// Each zone unit lies in one side-count run of one problem, so a block works
// on one problem (its dt, parameters and stop flag) at a time and writes one
// partial per unit; K4 runs one block per problem over that problem's units.

struct BatchArgs {
    FusedArgs a;                   // arena arrays; per-problem scalars come from probs
    const BUnit* units;
    int nunits;
    const PUnit* punits;
    int npunits;
    const BHeavy* heavy;
    int nheavy;
    const BProb* probs;
    Ctl* ctls;
    int rect_maps;                 // rect problems: points and corner slots computed
};

__device__ __forceinline__ void apply_prob(FusedArgs& a, const BProb& q) {
    a.gamma = q.gamma; a.ssmin = q.ssmin; a.alfa = q.alfa; a.qgamma = q.qgamma;
    a.q1 = q.q1; a.q2 = q.q2; a.cfl = q.cfl;
    a.nbcs = q.nbcs;
    a.bcv = q.bcvx;            // q.bcvy follows q.bcvx in memory
}
// A rect problem's zone kernels compute its zones' points and corner slots
// (offset to the problem's place in the arena) when `on`.
__device__ __forceinline__ void apply_rect(FusedArgs& a, const BProb& q, int on) {
    a.rect_nzx = on ? q.rect_nzx : 0;
    a.rect_m = q.rect_m;
    a.rect_sh = q.rect_sh;
    a.rect_so = q.so;
    a.rect_po = q.po;
}

// A batch unit's side work force slots: the unit's zones are a PN_BLOCK
// chunk of one side-count run, as side_w takes them.
__device__ __forceinline__ SideW unit_w(const BUnit& U, int z) {
    SideW w;
    w.u0 = U.sb;
    w.i = z - U.zb;
    w.n = U.ze - U.zb;
    return w;
}

template <int NS, int FORM, int CL>
__device__ __forceinline__ void pred_unit(const FusedArgs& a, const BUnit& U, double dt,
                                          double& dtmin, int& bad, ClCount& cnt) {
    const int z = U.zb + (int)threadIdx.x;
    const bool valid = z < U.ze;
    if constexpr (FORM == FORM_FP64) {
        if (valid) {
            const double d = zone_pred<NS>(a, z, U.sb + (z - U.zb) * NS, unit_w(U, z), dt, bad);
            dtmin = (d < dtmin) ? d : dtmin;
        }
    } else if constexpr (CL == CL_EXACT) {
        if (valid) {
            const double d = zone_pred_mixed<NS, CL_EXACT>(a, z, U.sb + (z - U.zb) * NS, unit_w(U, z), dt,
                                                           bad, true, cnt, a.ctl->zre_valid != 0);
            dtmin = (d < dtmin) ? d : dtmin;
        }
    } else {
        const int zc = valid ? z : U.ze - 1;
        const double d = zone_pred_mixed<NS, CL>(a, zc, U.sb + (zc - U.zb) * NS, unit_w(U, zc), dt, bad,
                                                 valid, cnt, a.ctl->zre_valid != 0);
        if (valid) dtmin = (d < dtmin) ? d : dtmin;
    }
}

template <int FORM, int CL>
__global__ void __launch_bounds__(PN_BLOCK)
__attribute__((amdgpu_waves_per_eu((CL == CL_LEARNED) ? 2 : ((FORM == FORM_MIXED && CL == CL_EXACT) ? 4 : 1))))
k_pred_batch(BatchArgs B) {
    // the learned form reads the network's weights from an LDS copy of the
    // block's own
    constexpr int NQ = (CL == CL_LEARNED) ? (int)(sizeof(ClosurePack) / 16) : 1;
    __shared__ uint4 wq[NQ];
    if constexpr (CL == CL_LEARNED) {
        const uint4* src = reinterpret_cast<const uint4*>(B.a.pack);
        for (int i = threadIdx.x; i < NQ; i += PN_BLOCK) wq[i] = src[i];
        __syncthreads();
    }
    // a unit's header (its work unit and its problem's done flag and dt) is
    // loaded one unit ahead, so a unit's zone loads do not wait behind it
    Units w = block_units(B.nunits, B.a.xcd);
    BUnit Un;
    int dn = 1;
    double dtn = 0.;
    if (w.u < w.end) {
        Un = B.units[w.u];
        dn = B.ctls[Un.pid].done;
        dtn = B.ctls[Un.pid].dt;
    }
    for (; w.u < w.end; w.u += w.step) {
        const BUnit U = Un;
        const int done = dn;
        const double dt = dtn;
        if (w.u + w.step < w.end) {
            Un = B.units[w.u + w.step];
            dn = B.ctls[Un.pid].done;
            dtn = B.ctls[Un.pid].dt;
        }
        if (done) continue;
        Ctl* c = B.ctls + U.pid;
        FusedArgs a = B.a;
        if constexpr (CL == CL_LEARNED) a.pack = reinterpret_cast<const ClosurePack*>(wq);
        apply_prob(a, B.probs[U.pid]);
        apply_rect(a, B.probs[U.pid], B.rect_maps);
        a.ctl = c;                 // the problem's zr/ze flag for the predictor
        double dtmin = 1.e99;
        int bad = 0;
        ClCount cnt;
        if (U.ns == 4) pred_unit<4, FORM, CL>(a, U, dt, dtmin, bad, cnt);
        else pred_unit<3, FORM, CL>(a, U, dt, dtmin, bad, cnt);
        const double m = block_min(dtmin);
        const int nb = block_sum(bad);
        if constexpr (CL == CL_LEARNED) {
            const int nl = block_sum(cnt.learned);
            const int nf = block_sum(cnt.fallback);
            if (threadIdx.x == 0 && (nl | nf)) {
                atomicAdd(a.cl_stats, (unsigned long long)nl);
                atomicAdd(a.cl_stats + 1, (unsigned long long)nf);
            }
        }
        if (threadIdx.x == 0) {
            a.part_dt[w.u] = m;
            a.part_bp[w.u] = nb;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(PN_BLOCK) k_point_batch(BatchArgs B) {
    if ((int)blockIdx.x < B.nheavy) {
        const BHeavy h = B.heavy[blockIdx.x];
        const Ctl* c = B.ctls + h.pid;
        if (c->done) return;
        FusedArgs a = B.a;
        apply_prob(a, B.probs[h.pid]);
        point_heavy<T>(a, h.p, c->dt);
        return;
    }
    const PUnit U = B.punits[blockIdx.x - B.nheavy];
    const Ctl* c = B.ctls + U.pid;
    if (c->done) return;
    FusedArgs a = B.a;
    apply_prob(a, B.probs[U.pid]);
    const int p = U.pb + (int)threadIdx.x;
    if (p < U.pe) point_light<T>(a, p, c->dt);
}

template <typename T>
__global__ void __launch_bounds__(PN_BLOCK) k_corr_batch(BatchArgs B) {
    // a unit's header (its work unit and its problem's done flag and dt) is
    // loaded one unit ahead, so a unit's zone loads do not wait behind it
    Units w = block_units(B.nunits, B.a.xcd);
    BUnit Un;
    int dn = 1;
    double dtn = 0.;
    if (w.u < w.end) {
        Un = B.units[w.u];
        dn = B.ctls[Un.pid].done;
        dtn = B.ctls[Un.pid].dt;
    }
    for (; w.u < w.end; w.u += w.step) {
        const BUnit U = Un;
        const int done = dn;
        const double dt = dtn;
        if (w.u + w.step < w.end) {
            Un = B.units[w.u + w.step];
            dn = B.ctls[Un.pid].done;
            dtn = B.ctls[Un.pid].dt;
        }
        if (done) continue;
        double dvov = 1.e-99;
        int bad = 0;
        const int z = U.zb + (int)threadIdx.x;
        if (z < U.ze) {
            if (U.ns == 4) {
                FusedArgs a = B.a;
                apply_rect(a, B.probs[U.pid], B.rect_maps);
                corr_zone<4, T>(a, z, U.sb + (z - U.zb) * 4, unit_w(U, z), dt, bad, dvov);
            } else {
                corr_zone<3, T>(B.a, z, U.sb + (z - U.zb) * 3, unit_w(U, z), dt, bad, dvov);
            }
        }
        const double m = block_max(dvov);
        const int nb = block_sum(bad);
        if (threadIdx.x == 0) {
            B.a.part_dv[w.u] = m;
            B.a.part_bc[w.u] = nb;
        }
    }
}

__global__ void __launch_bounds__(PN_FIN_BLOCK) k_finalize_batch(BatchArgs B, int zre) {
    const BProb& q = B.probs[blockIdx.x];
    const int n = q.uend - q.ubeg;
    finalize_block(B.ctls + blockIdx.x, B.a.part_dt + q.ubeg, B.a.part_bp + q.ubeg, n,
                   B.a.part_dv + q.ubeg, B.a.part_bc + q.ubeg, n, q.cflv, q.cstop, q.dtmax,
                   q.dtfac, q.tstop, zre);
}

// k_zre over the units of every problem whose flag is clear.
__global__ void __launch_bounds__(PN_BLOCK) k_zre_batch(BatchArgs B) {
    for (Units w = block_units(B.nunits, 0); w.u < w.end; w.u += w.step) {
        const BUnit U = B.units[w.u];
        if (B.ctls[U.pid].zre_valid) continue;
        const int z = U.zb + (int)threadIdx.x;
        if (z < U.ze) {
            const double zm = B.a.zm[z];
            B.a.zr[z] = zm / B.a.zvol[z];
            B.a.ze[z] = B.a.zetot[z] / (zm + 1.e-99);
        }
    }
}

BatchArgs make_batch_args(pennant_batch* b) {
    BatchArgs B;
    FusedArgs& a = B.a;
    std::memset(&a, 0, sizeof a);
    a.mapsp1 = b->mapsp1; a.pcbeg = b->pcbeg; a.pcpos = b->pcpos; a.smf = b->smf;
    a.nump = (int)b->nump;
    a.npown = a.nump;
    a.la_npred = 0;
    a.pxx = b->pxx; a.pxy = b->pxy; a.pux = b->pux; a.puy = b->puy;
    a.zr = b->zr; a.ze = b->ze; a.zetot = b->zetot; a.zwrate = b->zwrate; a.zvol = b->zvol;
    a.zp = b->zp; a.zm = b->zm;
    a.cfx = b->cfx; a.cfy = b->cfy; a.cmass = b->cmass; a.sfwx = b->sfwx; a.sfwy = b->sfwy;
    a.fcrn = b->fcrn; a.fsfw = b->fsfw;
    a.fsmf = b->fsmf;
    a.vs2 = b->vs2; a.pxpx = b->pxpx; a.pbc = b->pbc;
    a.part_dt = b->part_d; a.part_dv = b->part_d + b->nunits;
    a.part_bp = b->part_i; a.part_bc = b->part_i + b->nunits;
    a.xcd = b->tune.xcd;
    a.pack = (const ClosurePack*)b->closure;
    a.cap_buf = b->cap_buf;
    a.cap = b->cap;
    a.cap_count = b->cl_counters;
    a.cl_stats = b->cl_counters + 1;
    a.rect_nzx = 0;
    a.rect_m = 0;
    a.rect_sh = 0;
    a.rect_so = 0;
    a.rect_po = 0;
    B.rect_maps = b->tune.rect_maps;
    B.units = b->units; B.nunits = b->nunits;
    B.punits = b->punits; B.npunits = b->npunits;
    B.heavy = b->heavy; B.nheavy = b->nheavy;
    B.probs = b->probs; B.ctls = b->ctls;
    return B;
}

}  // namespace

namespace pn {

// Builds the zone runs, the per-point BC bits, the heavy-point list and the
// fused-path buffers (partials sized for the largest grids the tuning
// allows).
int refresh_side_fracs(pennant_problem* pb) {
    const int64_t S = pb->hm.nums;
    if (S > 0)
        hipLaunchKernelGGL(k_to_float, dim3(nblocks(S)), dim3(PN_BLOCK), 0, pb->stream, (int)S,
                           (const double*)pb->f("smf"), const_cast<float*>(pb->fsmf));
    pb->smf_stale = false;
    return (int)hipGetLastError();
}

// Whether the mesh is a structured rect in the generator's numbering, so
// that the fused kernels may compute its zones' points and corner slots
// (rect_zone) instead of loading mapsp1 and pcpos: one run of quads, sides
// 4z .. 4z + 3, and every side's point and slot equal to the formula's
// (checked here, with the same integer arithmetic as the device's).
void rect_detect(pennant_problem* pb, const std::vector<int>& pcpos) {
    const HostMesh& h = pb->hm;
    pb->rect_nzx = 0;
    const long nzx = pb->prm.nzx;
    if (pb->prm.meshtype != "rect" || own_points(pb) != h.nump || pb->zruns.size() != 3 ||
        pb->zruns[2] != 4 || nzx < 1 || (long)h.numz % nzx != 0 || (int64_t)h.nums != 4 * (int64_t)h.numz)
        return;
    int sh = 0;
    while ((1L << sh) < nzx) ++sh;
    const unsigned long long m = ((1ull << (32 + sh)) / (unsigned long long)nzx) + 1ull;
    const long npx = nzx + 1;
    for (int z = 0; z < h.numz; ++z) {
        const long zy = (long)(((unsigned long long)(unsigned)z * m) >> (32 + sh));
        const long zx = z - zy * nzx;
        if (zy != z / nzx || h.zsbeg[z] != 4 * z) return;
        const long p0 = z + zy;
        const int left = zx >= 1, below = zy >= 1;
        const long pts[4] = {p0, p0 + 1, p0 + 1 + npx, p0 + npx};
        const int rank[4] = {(left & below) + below + left, below + (below & (zx + 1 < nzx)), 0, left};
        for (int k = 0; k < 4; ++k)
            if (h.mapsp1[4 * z + k] != pts[k] || pcpos[4 * z + k] != (long)rank[k] * h.nump + pts[k]) return;
    }
    pb->rect_nzx = (int)nzx;
    pb->rect_m = m;
    pb->rect_sh = 32 + sh;
}

int fused_setup(pennant_problem* pb) {
    const HostMesh& h = pb->hm;
    pb->zruns.clear();
    for (int z = 0; z < h.numz;) {
        int e = z + 1;
        while (e < h.numz && h.znump[e] == h.znump[z]) ++e;
        pb->zruns.push_back(z);
        pb->zruns.push_back(e);
        pb->zruns.push_back(h.znump[z]);
        z = e;
    }
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
    std::vector<int> heavy;
    for (int p = 0; p < own_points(pb); ++p)
        if (h.pcbeg[p + 1] - h.pcbeg[p] > PN_HEAVY) heavy.push_back(p);
    auto al = [&](size_t n, size_t sz) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, n * sz > 0 ? n * sz : sz) != hipSuccess) return nullptr;
        if (hipMemset(p, 0, n * sz > 0 ? n * sz : sz) != hipSuccess) { (void)hipFree(p); return nullptr; }
        pb->allocs.push_back(p);
        return p;
    };
    const size_t S = h.nums, P = h.nump;
    if (!corner_slots_fit((int64_t)P, (int64_t)S)) return PENNANT_EINVAL;
    const size_t NC = (size_t)corner_slots((int64_t)P, (int64_t)S);
    pb->cfx = (double*)al(NC, 8); pb->cfy = (double*)al(NC, 8); pb->cmass = (double*)al(NC, 8);
    pb->sfwx = (double*)al(S, 8); pb->sfwy = (double*)al(S, 8);
    pb->vs2 = (double*)al(2 * P, 8); pb->pxpx = (double*)al(P, 8);
    pb->pbc = (unsigned char*)al(P, 1);
    pb->heavy = (int*)al(heavy.size(), 4);
    pb->nheavy = (int)heavy.size();
    // mixed-precision intermediates and the FP32 copy of smf
    pb->fcrn = (float*)al(3 * NC, 4);
    pb->fsfw = (float*)al(2 * S, 4);
    float* fsmf = (float*)al(S, 4);
    pb->fsmf = fsmf;
    if (!pb->fcrn || !pb->fsfw || !fsmf)
        return (int)hipErrorOutOfMemory;
    if (refresh_side_fracs(pb)) return (int)hipErrorUnknown;
    // K1 writes each corner's force and mass straight into its corner slot,
    // where K2 reads it.  (A piece's master point also lists the corners of
    // other pieces, pcside -1: their records arrive by the corner exchange.)
    std::vector<int> pcpos((size_t)h.nums);
    for (int p = 0; p < h.nump; ++p)
        for (int k = h.pcbeg[p]; k < h.pcbeg[p + 1]; ++k)
            if (h.pcside[k] >= 0) pcpos[h.pcside[k]] = (int)corner_slot((int64_t)P, p, k - h.pcbeg[p], k);
    pb->pcpos = (int*)al(S, 4);
    if (pb->pcpos &&
        hipMemcpy(pb->pcpos, pcpos.data(), S * 4, hipMemcpyHostToDevice) != hipSuccess)
        return (int)hipErrorUnknown;
    rect_detect(pb, pcpos);
    // partial slots: one per block of every K1 (resp. K3) launch of a cycle,
    // at the largest grid any tuning allows
    const int nruns = (int)pb->zruns.size() / 3;
    pb->npart_pred = nruns * PN_MAX_GRID;
    pb->npart_corr = nruns * PN_MAX_GRID;
    pb->cl_counters = (unsigned long long*)al(5, 8);
    pb->zpn = (double*)al((size_t)h.numz, 8);
    pb->zlist = (int*)al((size_t)h.numz, 4);
    pb->zcount = (unsigned*)al(pb->zruns.size() / 3 * PN_LSLICES * PN_LSTRIDE + 1, 4);
    if (!pb->cl_counters || !pb->zlist || !pb->zcount || !pb->zpn) return (int)hipErrorOutOfMemory;
    pb->part_d = (double*)al((size_t)pb->npart_pred + pb->npart_corr, 8);
    pb->part_i = (int*)al((size_t)pb->npart_pred + pb->npart_corr, 4);
    if (!pb->cfx || !pb->cfy || !pb->cmass || !pb->sfwx || !pb->sfwy || !pb->vs2 ||
        !pb->pxpx || !pb->pbc || !pb->heavy || !pb->pcpos || !pb->part_d || !pb->part_i)
        return (int)hipErrorOutOfMemory;
    if (!heavy.empty() &&
        hipMemcpy(pb->heavy, heavy.data(), heavy.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
        return (int)hipErrorUnknown;
    double bcv[16];
    for (int i = 0; i < 8; ++i) { bcv[i] = pb->bcvx[i]; bcv[8 + i] = pb->bcvy[i]; }
    pb->bcv_dev = (double*)al(16, 8);
    if (!pb->bcv_dev || hipMemcpy(pb->bcv_dev, bcv, sizeof bcv, hipMemcpyHostToDevice) != hipSuccess)
        return (int)hipErrorOutOfMemory;
    return (int)hipMemcpy(pb->pbc, bits.data(), bits.size(), hipMemcpyHostToDevice);
}

namespace {

// The learned pass over the zones a flagging predictor listed; with
// `pending` its counts go to the pending counters.
int launch_learned_pass(const pennant_problem* pb, const FusedArgs& a, bool pending, hipStream_t s) {
    FusedArgs b = a;
    if (pending) b.cl_stats = pb->cl_counters + 3;
    const HostMesh& h = pb->hm;
    for (size_t r = 0; r < pb->zruns.size(); r += 3) {
        const int zb = pb->zruns[r], ze = pb->zruns[r + 1], ns = pb->zruns[r + 2];
        const int sb = h.zsbeg[zb];
        // the learned pass holds 2 blocks per CU (registers): 512 blocks
        // fill the chip without dispatching blocks that find no work
        const int g = pred_grid(pb, ze - zb, ns, FORM_LEARNED);
        const dim3 grid(g < 512 ? g : 512), blk(PN_BLOCK);
        const int run = (int)(r / 3);
        switch (ns) {
            case 3: hipLaunchKernelGGL(k_pred_list<3>, grid, blk, 0, s, b, zb, ze, sb, run); break;
            case 4: hipLaunchKernelGGL(k_pred_list<4>, grid, blk, 0, s, b, zb, ze, sb, run); break;
            case 5: hipLaunchKernelGGL(k_pred_list<5>, grid, blk, 0, s, b, zb, ze, sb, run); break;
            case 6: hipLaunchKernelGGL(k_pred_list<6>, grid, blk, 0, s, b, zb, ze, sb, run); break;
            case 7: hipLaunchKernelGGL(k_pred_list<7>, grid, blk, 0, s, b, zb, ze, sb, run); break;
            case 8: hipLaunchKernelGGL(k_pred_list<8>, grid, blk, 0, s, b, zb, ze, sb, run); break;
            default: return PENNANT_EINVAL;
        }
    }
    return 0;
}

}  // namespace

int fused_cycle(pennant_problem* pb, hipEvent_t* ev, int form) {
    if (form == FORM_LEARNED && !pb->closure) return PENNANT_EINVAL;
    const bool mixed = (form != FORM_FP64);
    const FusedArgs a = make_args(pb);
    const Params& q = pb->prm;
    hipStream_t s = pb->stream;
    const HostMesh& h = pb->hm;
    if (ev) (void)hipEventRecord(ev[0], s);
    int poff = 0;
    for (size_t r = 0; r < pb->zruns.size(); r += 3) {
        const int zb = pb->zruns[r], ze = pb->zruns[r + 1], ns = pb->zruns[r + 2];
        const int sb = h.zsbeg[zb];
        switch (ns) {
            case 3: launch_pred<3>(pb, a, zb, ze, sb, poff, form, (int)(r / 3), s); break;
            case 4: launch_pred<4>(pb, a, zb, ze, sb, poff, form, (int)(r / 3), s); break;
            case 5: launch_pred<5>(pb, a, zb, ze, sb, poff, form, (int)(r / 3), s); break;
            case 6: launch_pred<6>(pb, a, zb, ze, sb, poff, form, (int)(r / 3), s); break;
            case 7: launch_pred<7>(pb, a, zb, ze, sb, poff, form, (int)(r / 3), s); break;
            case 8: launch_pred<8>(pb, a, zb, ze, sb, poff, form, (int)(r / 3), s); break;
            default: return PENNANT_EINVAL;
        }
        poff += pred_grid(pb, ze - zb, ns, form);
    }
    const int npred = poff;
    if (ev) (void)hipEventRecord(ev[1], s);
    if (form == FORM_LEARNED) {
        const int rc = launch_learned_pass(pb, a, false, s);
        if (rc) return rc;
    }
    if (ev) (void)hipEventRecord(ev[2], s);
    const dim3 pgrid(pb->nheavy + nblocks(own_points(pb)));
    if (mixed)
        hipLaunchKernelGGL(k_point<float>, pgrid, dim3(PN_BLOCK), 0, s, a);
    else
        hipLaunchKernelGGL(k_point<double>, pgrid, dim3(PN_BLOCK), 0, s, a);
    if (ev) (void)hipEventRecord(ev[3], s);
    poff = 0;
    for (size_t r = 0; r < pb->zruns.size(); r += 3) {
        const int zb = pb->zruns[r], ze = pb->zruns[r + 1], ns = pb->zruns[r + 2];
        const int sb = h.zsbeg[zb];
        switch (ns) {
            case 3: launch_corr<3>(pb, a, zb, ze, sb, poff, form, s); break;
            case 4: launch_corr<4>(pb, a, zb, ze, sb, poff, form, s); break;
            case 5: launch_corr<5>(pb, a, zb, ze, sb, poff, form, s); break;
            case 6: launch_corr<6>(pb, a, zb, ze, sb, poff, form, s); break;
            case 7: launch_corr<7>(pb, a, zb, ze, sb, poff, form, s); break;
            case 8: launch_corr<8>(pb, a, zb, ze, sb, poff, form, s); break;
            default: return PENNANT_EINVAL;
        }
        poff += corr_grid(pb, ze - zb);
    }
    const int ncorr = poff;
    if (ev) (void)hipEventRecord(ev[4], s);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(PN_FIN_BLOCK), 0, s, a, npred, ncorr, q.cflv,
                       q.cstop, q.dtmax, q.dtfac, q.tstop, mixed ? 0 : 1);
    if (ev) (void)hipEventRecord(ev[5], s);
    return (int)hipGetLastError();
}

namespace {

DtParams dt_params(const pennant_problem* pb) {
    const Params& q = pb->prm;
    DtParams d;
    d.cflv = q.cflv; d.dtmax = q.dtmax; d.dtfac = q.dtfac; d.tstop = q.tstop; d.cstop = q.cstop;
    return d;
}

// Predictor launches of the pipelined cycle (the prologue's, or with REDO
// the redo), writing zp to zpn; returns the number of partials.
template <bool REDO>
int launch_lookahead_pred(const pennant_problem* pb, const FusedArgs& a, int form, hipStream_t s) {
    FusedArgs b = a;
    b.zp = pb->zpn;
    const HostMesh& h = pb->hm;
    int poff = 0;
    for (size_t r = 0; r < pb->zruns.size(); r += 3) {
        const int zb = pb->zruns[r], ze = pb->zruns[r + 1], ns = pb->zruns[r + 2];
        const int sb = h.zsbeg[zb];
        const dim3 grid(pred_grid(pb, ze - zb, ns, form)), blk(PN_BLOCK);
        const int run = (int)(r / 3);
        const bool fl = (form == FORM_LEARNED);
#define PN_LA(NS)                                                                                  \
    if (fl) hipLaunchKernelGGL((k_pred_mixed<NS, true, REDO>), grid, blk, 0, s, b, zb, ze, sb, poff, run); \
    else hipLaunchKernelGGL((k_pred_mixed<NS, false, REDO>), grid, blk, 0, s, b, zb, ze, sb, poff, run);
        switch (ns) {
            case 3: PN_LA(3) break;
            case 4: PN_LA(4) break;
            case 5: PN_LA(5) break;
            case 6: PN_LA(6) break;
            case 7: PN_LA(7) break;
            case 8: PN_LA(8) break;
            default: return -1;
        }
#undef PN_LA
        poff += grid.x;
    }
    return poff;
}


}  // namespace

int fused_prologue(pennant_problem* pb, int form) {
    if (form == FORM_FP64 || (form == FORM_LEARNED && !pb->closure)) return PENNANT_EINVAL;
    const FusedArgs a = make_args(pb);
    hipStream_t s = pb->stream;
    const int nz = (int)pb->zruns.size() / 3;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, pb->zcount, nz * PN_LSLICES * PN_LSTRIDE);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, pb->cl_counters + 3, 2);
    const int npred = launch_lookahead_pred<false>(pb, a, form, s);
    if (npred < 0) return PENNANT_EINVAL;
    hipLaunchKernelGGL(k_finalize_la<false>, dim3(1), dim3(PN_FIN_BLOCK), 0, s, a, npred);
    if (form == FORM_LEARNED) {
        const int rc = launch_learned_pass(pb, a, true, s);
        if (rc) return rc;
    }
    return (int)hipGetLastError();
}

int fused_cycle_pipe(pennant_problem* pb, hipEvent_t* ev, int form) {
    if (form == FORM_FP64 || (form == FORM_LEARNED && !pb->closure)) return PENNANT_EINVAL;
    const FusedArgs a = make_args(pb);
    const DtParams q = dt_params(pb);
    hipStream_t s = pb->stream;
    const HostMesh& h = pb->hm;
    const bool fl = (form == FORM_LEARNED);
    // partials of one predictor pass (the zone pass's, or its redo's)
    int nparts = 0;
    for (size_t r = 0; r < pb->zruns.size(); r += 3)
        nparts += pred_grid(pb, pb->zruns[r + 1] - pb->zruns[r], pb->zruns[r + 2], form);
    if (ev) (void)hipEventRecord(ev[0], s);
    {
        FusedArgs a2 = a;
        a2.la_npred = nparts;   // and one block more, for the look-ahead fold
        hipLaunchKernelGGL(k_point<float>, dim3(pb->nheavy + nblocks(own_points(pb)) + 1), dim3(PN_BLOCK), 0, s,
                           a2);
    }
    if (ev) (void)hipEventRecord(ev[1], s);
    int poff = 0;
    for (size_t r = 0; r < pb->zruns.size(); r += 3) {
        const int zb = pb->zruns[r], ze = pb->zruns[r + 1], ns = pb->zruns[r + 2];
        const int sb = h.zsbeg[zb];
        const dim3 grid(pred_grid(pb, ze - zb, ns, form)), blk(PN_BLOCK);
        const int run = (int)(r / 3);
#define PN_ZP(NS)                                                                                  \
    if (fl) hipLaunchKernelGGL((k_zone_pipe<NS, true>), grid, blk, 0, s, a, zb, ze, sb, poff, run, q); \
    else hipLaunchKernelGGL((k_zone_pipe<NS, false>), grid, blk, 0, s, a, zb, ze, sb, poff, run, q);
        switch (ns) {
            case 3: PN_ZP(3) break;
            case 4: PN_ZP(4) break;
            case 5: PN_ZP(5) break;
            case 6: PN_ZP(6) break;
            case 7: PN_ZP(7) break;
            case 8: PN_ZP(8) break;
            default: return PENNANT_EINVAL;
        }
#undef PN_ZP
        poff += grid.x;
    }
    if (ev) (void)hipEventRecord(ev[2], s);
    hipLaunchKernelGGL(k_finalize_pipe, dim3(1), dim3(PN_FIN_BLOCK), 0, s, a, poff, poff, q);
    // the redo, when the look-ahead dt was not the cycle's dt; its partials
    // are folded by the next cycle's K2 (every launch sequence that does not
    // continue with this form's cycles starts again from the prologue)
    const int npred = launch_lookahead_pred<true>(pb, a, form, s);
    if (npred != poff || npred != nparts) return PENNANT_EINVAL;
    if (ev) (void)hipEventRecord(ev[3], s);
    if (fl) {
        const int rc = launch_learned_pass(pb, a, true, s);
        if (rc) return rc;
    }
    if (ev) (void)hipEventRecord(ev[4], s);
    return (int)hipGetLastError();
}

int reset_zone_lists(pennant_problem* pb) {
    const int nz = (int)pb->zruns.size() / 3;
    if (pb->zcount && nz > 0)
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, pb->stream, pb->zcount, nz * PN_LSLICES * PN_LSTRIDE);
    return (int)hipGetLastError();
}

int materialize_zre(pennant_problem* pb) {
    if (!pb->zre_stale) return 0;
    const FusedArgs a = make_args(pb);
    const int nz = pb->hm.numz;
    hipLaunchKernelGGL(k_zre, dim3(nblocks(nz)), dim3(PN_BLOCK), 0, pb->stream, a, nz);
    hipLaunchKernelGGL(k_zre_done, dim3(1), dim3(64), 0, pb->stream, pb->ctl, 1);
    pb->zre_stale = false;
    return (int)hipGetLastError();
}

int batch_materialize_zre(pennant_batch* b) {
    if (!b->zre_stale) return 0;
    const BatchArgs B = make_batch_args(b);
    hipLaunchKernelGGL(k_zre_batch, dim3(grid_for(b->nunits, PN_MAX_GRID)), dim3(PN_BLOCK), 0,
                       b->stream, B);
    hipLaunchKernelGGL(k_zre_done, dim3(1), dim3(64), 0, b->stream, b->ctls, b->nprob);
    b->zre_stale = false;
    return (int)hipGetLastError();
}

int launch_to_float(const double* src, float* dst, int64_t n, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_to_float, dim3(nblocks(n)), dim3(PN_BLOCK), 0, s, (int)n, src, dst);
    return (int)hipGetLastError();
}

int batch_cycle(pennant_batch* b, int form) {
    if (form == FORM_LEARNED && !b->closure) return PENNANT_EINVAL;
    const BatchArgs B = make_batch_args(b);
    hipStream_t s = b->stream;
    const dim3 zg(grid_for(b->nunits, b->tune.max_blocks)), blk(PN_BLOCK);
    if (form == FORM_FP64)
        hipLaunchKernelGGL((k_pred_batch<FORM_FP64, CL_EXACT>), zg, blk, 0, s, B);
    else if (form == FORM_LEARNED)
        hipLaunchKernelGGL((k_pred_batch<FORM_MIXED, CL_LEARNED>), zg, blk, 0, s, B);
    else if (b->cap_buf)
        hipLaunchKernelGGL((k_pred_batch<FORM_MIXED, CL_CAPTURE>), zg, blk, 0, s, B);
    else
        hipLaunchKernelGGL((k_pred_batch<FORM_MIXED, CL_EXACT>), zg, blk, 0, s, B);
    const dim3 pg(b->nheavy + b->npunits);
    if (form == FORM_FP64) {
        hipLaunchKernelGGL(k_point_batch<double>, pg, blk, 0, s, B);
        hipLaunchKernelGGL(k_corr_batch<double>, zg, blk, 0, s, B);
    } else {
        hipLaunchKernelGGL(k_point_batch<float>, pg, blk, 0, s, B);
        hipLaunchKernelGGL(k_corr_batch<float>, zg, blk, 0, s, B);
    }
    hipLaunchKernelGGL(k_finalize_batch, dim3(b->nprob), dim3(PN_FIN_BLOCK), 0, s, B,
                       form == FORM_FP64 ? 1 : 0);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Several pieces of one mesh on one device (src/Mesh.cc:432-617 partitions
// the points into private, master and ghost sets; src/Hydro.cc:450-455,
// 571-574, 766-769, 866-867, 902-906 read ghosts and reduce shared points).
// Each piece runs the fused K1 / K2 / K3 over its own arrays.  Between them:
//   corner exchange  each corner a piece's zones have at a ghost point is
//                    copied from the slot the piece's K1 wrote to the slot of
//                    the owner's corner list that holds it; the owner lists a
//                    master point's corners from every piece in ascending
//                    one-piece side order, so K2's ordered sum is the
//                    one-piece run's sum, bit for bit;
//   halo             after K2 the owners' new point state (px, pu and the
//                    calcWork point values) goes to the ghost copies, which
//                    K3 and the next cycle's K1 read.
// The pieces share one control block and one array of block partials, so a
// single finalize folds every piece's dt, dvol and bad-side partials.

namespace {

template <typename T>
__global__ void __launch_bounds__(PN_BLOCK) k_corner_xchg(const PieceTab* tab, const XEnt* x, int n,
                                                          const Ctl* ctl) {
    if (ctl->done) return;
    const int i = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (i >= n) return;
    const XEnt e = x[i];
    const PieceTab& a = tab[e.sp];
    const PieceTab& b = tab[e.dp];
    if constexpr (sizeof(T) == 4) {
        const int64_t u = 3 * (int64_t)e.src, v = 3 * (int64_t)e.dst;
        b.fcrn[v] = a.fcrn[u];
        b.fcrn[v + 1] = a.fcrn[u + 1];
        b.fcrn[v + 2] = a.fcrn[u + 2];
    } else {
        b.cfx[e.dst] = a.cfx[e.src];
        b.cfy[e.dst] = a.cfy[e.src];
        b.cmass[e.dst] = a.cmass[e.src];
    }
}

template <typename T>
__global__ void __launch_bounds__(PN_BLOCK) k_halo(const PieceTab* tab, const XEnt* x, int n,
                                                   const Ctl* ctl) {
    if (ctl->done) return;
    const int i = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (i >= n) return;
    const XEnt e = x[i];
    const PieceTab& a = tab[e.sp];
    const PieceTab& b = tab[e.dp];
    b.pxx[e.dst] = a.pxx[e.src];
    b.pxy[e.dst] = a.pxy[e.src];
    b.pux[e.dst] = a.pux[e.src];
    b.puy[e.dst] = a.puy[e.src];
    b.vs2[2 * (int64_t)e.dst] = a.vs2[2 * (int64_t)e.src];
    b.vs2[2 * (int64_t)e.dst + 1] = a.vs2[2 * (int64_t)e.src + 1];
    if constexpr (sizeof(T) == 8) b.pxpx[e.dst] = a.pxpx[e.src];
}

FusedArgs piece_args(pennant_pieces* pp, int q) {
    FusedArgs a = make_args(pp->pc[q]);
    a.ctl = pp->pc[0]->ctl;
    a.part_dt = pp->part_d;
    a.part_dv = pp->part_d + pp->npart;
    a.part_bp = pp->part_i;
    a.part_bc = pp->part_i + pp->npart;
    // no learned pass or tuple capture in a multi-piece run; the pieces'
    // maps are loaded
    a.rect_nzx = 0;
    a.zlist = nullptr;
    a.zcount = nullptr;
    a.nzruns = 0;
    a.cap_buf = nullptr;
    return a;
}

}  // namespace

// Phase boundaries of the piece cycle: every piece's stream waits for piece
// 0's stream (fork), and piece 0's stream waits for every piece (join).
namespace {
int pieces_fork(pennant_pieces* pp, int phase) {
    hipStream_t s0 = pp->pc[0]->stream;
    int rc = (int)hipEventRecord(pp->ev_fork[(size_t)phase], s0);
    for (size_t q = 1; q < pp->pc.size() && !rc; ++q)
        rc = (int)hipStreamWaitEvent(pp->pc[q]->stream, pp->ev_fork[(size_t)phase], 0);
    return rc;
}
int pieces_join(pennant_pieces* pp, int phase) {
    hipStream_t s0 = pp->pc[0]->stream;
    const size_t n = pp->pc.size();
    int rc = 0;
    for (size_t q = 1; q < n && !rc; ++q) {
        hipEvent_t e = pp->ev_join[(size_t)phase * n + q];
        rc = (int)hipEventRecord(e, pp->pc[q]->stream);
        if (!rc) rc = (int)hipStreamWaitEvent(s0, e, 0);
    }
    return rc;
}
}  // namespace

int pieces_cycle(pennant_pieces* pp, int form) {
    if (form == FORM_LEARNED || pp->pc.empty()) return PENNANT_EINVAL;
    const bool mixed = (form != FORM_FP64);
    const int n = (int)pp->pc.size();
    hipStream_t s0 = pp->pc[0]->stream;
    std::vector<FusedArgs> A((size_t)n);
    for (int q = 0; q < n; ++q) A[(size_t)q] = piece_args(pp, q);
    int rc = pieces_fork(pp, 0);
    if (rc) return rc;
    int poff = 0;
    for (int q = 0; q < n; ++q) {
        const pennant_problem* pb = pp->pc[(size_t)q];
        hipStream_t s = pb->stream;
        for (size_t r = 0; r < pb->zruns.size(); r += 3) {
            const int zb = pb->zruns[r], ze = pb->zruns[r + 1], ns = pb->zruns[r + 2];
            const int sb = pb->hm.zsbeg[zb];
            const int run = (int)(r / 3);
            if (poff + pred_grid(pb, ze - zb, ns, form) > pp->npart) return PENNANT_EINVAL;
            switch (ns) {
                case 3: launch_pred<3>(pb, A[(size_t)q], zb, ze, sb, poff, form, run, s); break;
                case 4: launch_pred<4>(pb, A[(size_t)q], zb, ze, sb, poff, form, run, s); break;
                case 5: launch_pred<5>(pb, A[(size_t)q], zb, ze, sb, poff, form, run, s); break;
                case 6: launch_pred<6>(pb, A[(size_t)q], zb, ze, sb, poff, form, run, s); break;
                case 7: launch_pred<7>(pb, A[(size_t)q], zb, ze, sb, poff, form, run, s); break;
                case 8: launch_pred<8>(pb, A[(size_t)q], zb, ze, sb, poff, form, run, s); break;
                default: return PENNANT_EINVAL;
            }
            poff += pred_grid(pb, ze - zb, ns, form);
        }
    }
    const int npred = poff;
    if ((rc = pieces_join(pp, 0))) return rc;
    const Ctl* ctl = pp->pc[0]->ctl;
    if (pp->ncx > 0) {
        if (mixed)
            hipLaunchKernelGGL(k_corner_xchg<float>, dim3(nblocks(pp->ncx)), dim3(PN_BLOCK), 0, s0, pp->tab,
                               pp->cx, pp->ncx, ctl);
        else
            hipLaunchKernelGGL(k_corner_xchg<double>, dim3(nblocks(pp->ncx)), dim3(PN_BLOCK), 0, s0, pp->tab,
                               pp->cx, pp->ncx, ctl);
    }
    if ((rc = pieces_fork(pp, 1))) return rc;
    for (int q = 0; q < n; ++q) {
        const pennant_problem* pb = pp->pc[(size_t)q];
        const dim3 pgrid(pb->nheavy + nblocks(own_points(pb)));
        if (mixed)
            hipLaunchKernelGGL(k_point<float>, pgrid, dim3(PN_BLOCK), 0, pb->stream, A[(size_t)q]);
        else
            hipLaunchKernelGGL(k_point<double>, pgrid, dim3(PN_BLOCK), 0, pb->stream, A[(size_t)q]);
    }
    if ((rc = pieces_join(pp, 1))) return rc;
    if (pp->nhalo > 0) {
        if (mixed)
            hipLaunchKernelGGL(k_halo<float>, dim3(nblocks(pp->nhalo)), dim3(PN_BLOCK), 0, s0, pp->tab, pp->halo,
                               pp->nhalo, ctl);
        else
            hipLaunchKernelGGL(k_halo<double>, dim3(nblocks(pp->nhalo)), dim3(PN_BLOCK), 0, s0, pp->tab, pp->halo,
                               pp->nhalo, ctl);
    }
    if ((rc = pieces_fork(pp, 2))) return rc;
    poff = 0;
    for (int q = 0; q < n; ++q) {
        const pennant_problem* pb = pp->pc[(size_t)q];
        hipStream_t s = pb->stream;
        for (size_t r = 0; r < pb->zruns.size(); r += 3) {
            const int zb = pb->zruns[r], ze = pb->zruns[r + 1], ns = pb->zruns[r + 2];
            const int sb = pb->hm.zsbeg[zb];
            if (poff + corr_grid(pb, ze - zb) > pp->npart) return PENNANT_EINVAL;
            switch (ns) {
                case 3: launch_corr<3>(pb, A[(size_t)q], zb, ze, sb, poff, form, s); break;
                case 4: launch_corr<4>(pb, A[(size_t)q], zb, ze, sb, poff, form, s); break;
                case 5: launch_corr<5>(pb, A[(size_t)q], zb, ze, sb, poff, form, s); break;
                case 6: launch_corr<6>(pb, A[(size_t)q], zb, ze, sb, poff, form, s); break;
                case 7: launch_corr<7>(pb, A[(size_t)q], zb, ze, sb, poff, form, s); break;
                case 8: launch_corr<8>(pb, A[(size_t)q], zb, ze, sb, poff, form, s); break;
                default: return PENNANT_EINVAL;
            }
            poff += corr_grid(pb, ze - zb);
        }
    }
    const int ncorr = poff;
    if ((rc = pieces_join(pp, 2))) return rc;
    const Params& q = pp->prm;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(PN_FIN_BLOCK), 0, s0, A[0], npred, ncorr, q.cflv, q.cstop,
                       q.dtmax, q.dtfac, q.tstop, mixed ? 0 : 1);
    if (mixed) pp->zre_stale = true;
    return (int)hipGetLastError();
}

int pieces_materialize_zre(pennant_pieces* pp) {
    if (!pp->zre_stale || pp->pc.empty()) return 0;
    hipStream_t s = pp->pc[0]->stream;
    for (size_t q = 0; q < pp->pc.size(); ++q) {
        const int nz = pp->pc[q]->hm.numz;
        if (nz > 0)
            hipLaunchKernelGGL(k_zre, dim3(nblocks(nz)), dim3(PN_BLOCK), 0, s, piece_args(pp, (int)q), nz);
    }
    hipLaunchKernelGGL(k_zre_done, dim3(1), dim3(64), 0, s, pp->pc[0]->ctl, 1);
    pp->zre_stale = false;
    return (int)hipGetLastError();
}

}  // namespace pn
