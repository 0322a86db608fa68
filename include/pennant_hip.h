/*
 * pennant_hip.h -- C ABI of the MI355X (gfx950) PENNANT hydro path.
 *
 * Two layers:
 *
 *  1. Task entry points, pennant_hip_<task>(): one per Legion task ID on the
 *     per-cycle hydro path.  Each replaces the TOC_PROC (CUDA) variant of that
 *     task ID (see the citation on each declaration) and takes exactly the
 *     fields of the task's region requirements, as device pointers, in the
 *     reference's positional order.  A Legion HIP task wrapper unpacks its
 *     accessors with acc.ptr(rect) and forwards them.
 *
 *  2. Driver entry points, pennant_*(): a whole-problem driver (deck parsing,
 *     mesh generation, initialisation, the cycle loop with dt kept on the
 *     device) mirroring Driver/Hydro (src/Driver.cc, src/Hydro.cc).  It runs
 *     the cycle either through the task entry points (one launch per task, the
 *     reference's structure) or through the fused cycle kernels.
 *
 * Conventions for every task entry point:
 *   - all array arguments are device pointers; a vector (double2) field is a
 *     pennant_vec / pennant_cvec {x, y, stride}: element i is (x[i*stride],
 *     y[i*stride]).  The reference's double2 field instance
 *     (AccessorRO<double2>::ptr(rect), src/QCS.cc:135-141, src/Mesh.cu:57-62)
 *     is passed as is: x = (double*)ptr, y = x + 1, stride = 2.  Two separate
 *     arrays take stride 1 (stride 0 means 1, so {x, y} initialisers keep
 *     working).  Interleaved fields with 16-byte aligned x take 16-byte
 *     accesses; any other stride >= 1 works through a strided path;
 *   - scalar arguments that the reference passes in task->args are passed by
 *     value; dt (task->futures[0] in the reference) is a device pointer to one
 *     double, so a captured hipGraph can replay with a new dt;
 *   - launches go on `stream`; nothing is allocated, nothing synchronises the
 *     host; every result is deterministic (zone sums run over the zone's sides
 *     in ascending side order and point sums over a point-corner CSR sorted by
 *     side, which reproduces the reference's serial summation order);
 *   - return value: 0, or a hipError_t code (or PENNANT_EINVAL for arguments
 *     the kernels cannot accept).  Nothing calls exit().
 */
#ifndef PENNANT_HIP_H_
#define PENNANT_HIP_H_

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PENNANT_EINVAL (-22)
#define PENNANT_EBADSIDES (-100)

/* Mesh connectivity on the device.  Counts follow src/GenMesh.cc:469-526;
 * maps follow src/Mesh.hh:33-105 (MAPSP1, MAPSP2, MAPSZ, MAPSS3, MAPSS4,
 * ZNUMP) stored as int32.  Sides of a zone are contiguous
 * (zsbeg[z]..zsbeg[z+1]-1) and cyclic, as the reference generates them
 * (src/GenMesh.cc:1007-1020,1207-1249): mapss3[s] / mapss4[s] are the
 * previous / next side of the zone (s-1 and s+1, wrapping at the zone's
 * ends) and mapsp2[s] == mapsp1[mapss4[s]].  Zones have 3 to
 * PENNANT_MAX_ZSIDES sides; maxznump is the largest.  pcbeg/pcside is the
 * point-corner CSR: for point p, pcside[pcbeg[p]..pcbeg[p+1]-1] lists, in
 * ascending order, the sides s with mapsp1[s] == p (corner s of the
 * reference, src/Hydro.cc:1100-1116). */
typedef struct pennant_mesh {
    int32_t nump, numz, nums;
    int32_t maxznump;          /* largest znump; at most PENNANT_MAX_ZSIDES */
    const int32_t* mapsp1;
    const int32_t* mapsp2;
    const int32_t* mapsz;
    const int32_t* mapss3;
    const int32_t* mapss4;
    const int32_t* znump;
    const int32_t* zsbeg;      /* numz + 1 entries */
    const int32_t* pcbeg;      /* nump + 1 entries */
    const int32_t* pcside;     /* nums entries */
} pennant_mesh;

#define PENNANT_MAX_ZSIDES 8

typedef struct { double* x; double* y; int64_t stride; } pennant_vec;
typedef struct { const double* x; const double* y; int64_t stride; } pennant_cvec;

/* Bytes of device workspace the reduction tasks (calc_vols, calc_dt_new,
 * calc_dvol) need for their per-block partials. */
size_t pennant_hip_reduce_workspace_bytes(void);

/* --- Mesh (src/Mesh.hh:107-122 task IDs) ------------------------------- */

/* TID_CALCCTRS: ex = (px1+px2)/2 per side; zx = sum px1*(1/n) per zone.
 * Replaces gpu_calc_ctrs, src/Mesh.cu:45-115 (CPU body src/Mesh.cc:797-837). */
int pennant_hip_calc_ctrs(const pennant_mesh* m, pennant_cvec px,
                          pennant_vec ex, pennant_vec zx, hipStream_t stream);

/* TID_CALCVOLS: sarea, svol per side; zarea, zvol per zone; *numsbad (device
 * int) receives the count of sides with svol <= 0 (the task's SumOp<int>
 * future; an integer sum, exact in any order).  `workspace` is not used and
 * may be NULL.  Replaces gpu_calc_vols, src/Mesh.cu:117-210
 * (src/Mesh.cc:928-989). */
int pennant_hip_calc_vols(const pennant_mesh* m, pennant_cvec px, pennant_cvec zx,
                          double* sarea, double* svol, double* zarea, double* zvol,
                          int32_t* numsbad, void* workspace, hipStream_t stream);

/* TID_CALCSIDEFRACS (init only): smf = sarea / zarea[z].
 * src/Mesh.cc:1085-1101. */
int pennant_hip_calc_side_fracs(const pennant_mesh* m, const double* sarea,
                                const double* zarea, double* smf, hipStream_t stream);

/* TID_CALCSURFVECS: ssurf = rotateCCW(ex - zx[z]).
 * Replaces src/Mesh.cu:266-305 (src/Mesh.cc:1104-1123). */
int pennant_hip_calc_surf_vecs(const pennant_mesh* m, pennant_cvec ex, pennant_cvec zx,
                               pennant_vec ssurf, hipStream_t stream);

/* TID_CALCEDGELEN: elen = |px[mapsp2] - px[mapsp1]| per side (reads mapsp1
 * and mapsp2; EINVAL if either is NULL).
 * Replaces src/Mesh.cu:212-264 (src/Mesh.cc:1151-1179). */
int pennant_hip_calc_edge_len(const pennant_mesh* m, pennant_cvec px, double* elen,
                              hipStream_t stream);

/* TID_CALCCHARLEN: zdl = min(1e99, min_s fac*sarea/elen), fac = 3 for
 * triangles else 4.  Replaces src/Mesh.cu:307-372 (src/Mesh.cc:1216-1244). */
int pennant_hip_calc_char_len(const pennant_mesh* m, const double* sarea,
                              const double* elen, double* zdl, hipStream_t stream);

/* --- Hydro (src/Hydro.hh:35-51 task IDs) ------------------------------- */

/* TID_ADVPOSHALF: pxp = px0 + (dt/2)*pu0.
 * Replaces gpu_adv_pos_half, src/Hydro.cu:82-122 (src/Hydro.cc:986-1006). */
int pennant_hip_adv_pos_half(const pennant_mesh* m, pennant_cvec px0, pennant_cvec pu0,
                             pennant_vec pxp, const double* dt, hipStream_t stream);

/* TID_CALCRHO: zr = zm / zvol.
 * Replaces src/Hydro.cu:124-163 (src/Hydro.cc:1034-1054). */
int pennant_hip_calc_rho(const pennant_mesh* m, const double* zm, const double* zvol,
                         double* zr, hipStream_t stream);

/* TID_CALCCRNRMASS: pmaswt[p] += zr*zarea*0.5*(smf[s]+smf[ss3[s]]) over the
 * point's corners (ascending s).  Replaces gpu_calc_crnr_mass,
 * src/Hydro.cu:165-224 (src/Hydro.cc:1083-1117). */
int pennant_hip_calc_crnr_mass(const pennant_mesh* m, const double* smf,
                               const double* zr, const double* zarea,
                               double* pmaswt, hipStream_t stream);

/* TID_SUMCRNRFORCE: pf[p] += (sfp+sfq+sft)[s] - (sfp+sfq+sft)[ss3[s]] over
 * the point's corners (ascending s).  Replaces gpu_sum_crnr_force,
 * src/Hydro.cu:226-283 (src/Hydro.cc:1159-1193). */
int pennant_hip_sum_crnr_force(const pennant_mesh* m, pennant_cvec sfp, pennant_cvec sfq,
                               pennant_cvec sft, pennant_vec pf, hipStream_t stream);

/* TID_CALCACCEL: pa = pf / max(pmaswt, 1e-99).
 * Replaces src/Hydro.cu:285-323 (src/Hydro.cc:1235-1254). */
int pennant_hip_calc_accel(const pennant_mesh* m, pennant_cvec pf, const double* pmaswt,
                           pennant_vec pa, hipStream_t stream);

/* TID_ADVPOSFULL: pu = pu0 + dt*pa; px = px0 + (dt*0.5)*(pu0+pu).
 * Replaces src/Hydro.cu:325-371 (src/Hydro.cc:1281-1306). */
int pennant_hip_adv_pos_full(const pennant_mesh* m, pennant_cvec px0, pennant_cvec pu0,
                             pennant_cvec pa, pennant_vec px, pennant_vec pu,
                             const double* dt, hipStream_t stream);

/* TID_CALCWORK: zw = sum dwork (from 0), zetot += dwork per side (ascending),
 * dwork = -(dt/2)*(dot(sfp+sfq, pu0[p1]+pu[p1])*pxp[p1].x
 *                  + dot(-(sfp+sfq), pu0[p2]+pu[p2])*pxp[p2].x).
 * Replaces gpu_calc_work, src/Hydro.cu:373-473 (src/Hydro.cc:1339-1406). */
int pennant_hip_calc_work(const pennant_mesh* m, pennant_cvec sfp, pennant_cvec sfq,
                          pennant_cvec pu, pennant_cvec pu0, pennant_cvec pxp,
                          double* zw, double* zetot, const double* dt, hipStream_t stream);

/* TID_CALCWORKRATE: zwrate = (zw + zp*(zvol-zvol0)) * (1/dt).
 * Replaces src/Hydro.cu:475-522 (src/Hydro.cc:1483-1510). */
int pennant_hip_calc_work_rate(const pennant_mesh* m, const double* zvol0,
                               const double* zvol, const double* zw, const double* zp,
                               double* zwrate, const double* dt, hipStream_t stream);

/* TID_CALCENERGY: ze = zetot / (zm + 1e-99).
 * Replaces src/Hydro.cu:524-561 (src/Hydro.cc:1545-1564). */
int pennant_hip_calc_energy(const pennant_mesh* m, const double* zetot, const double* zm,
                            double* ze, hipStream_t stream);

/* TID_CALCDTNEW: *dtnew = min(1e99, min_z zdl*cfl/max(zdu, max(zss, 1e-99))).
 * Replaces gpu_calc_dt_new + reduce_double<MinOp>, src/Hydro.cu:563-626 and
 * src/CudaHelp.hh:14-41 (src/Hydro.cc:1591-1619). */
int pennant_hip_calc_dt_new(const pennant_mesh* m, const double* zdl, const double* zdu,
                            const double* zss, double cfl, double* dtnew,
                            void* workspace, hipStream_t stream);

/* pennant_hip_calc_dt_new without a workspace argument.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same kernels and bits as pennant_hip_calc_dt_new; the
 * per-block partials live in a library-owned device array of PN_RED_BLOCKS
 * (1024) doubles per device, which the bounded reduction grid never exceeds
 * for any numz.  Allocates nothing.  Calls that overlap on one device race on
 * that array: the caller orders them.  PENNANT_EINVAL for a NULL mesh or
 * output, or numz < 0. */
int pennant_hip_calc_dt_new_noalloc(const pennant_mesh* m, const double* zdl, const double* zdu,
                                    const double* zss, double cfl, double* dtnew,
                                    hipStream_t stream);

/* TID_CALCDVOL: *dvovmax = max(1e-99, max_z |(zvol-zvol0)/zvol0|).
 * Replaces gpu_calc_dvol, src/Hydro.cu:628-681 (src/Hydro.cc:1622-1644). */
int pennant_hip_calc_dvol(const pennant_mesh* m, const double* zvol, const double* zvol0,
                          double* dvovmax, void* workspace, hipStream_t stream);

/* TID_CALCDT: *dtrec = min(1e99, *dtnew, *dtlast*cflv / *dvovmax) on the
 * device (a one-thread kernel; the reference runs it as a host task,
 * src/Hydro.cc:1647-1667). */
int pennant_hip_calc_dt(const double* dtlast, const double* dtnew, const double* dvovmax,
                        double cflv, double* dtrec, hipStream_t stream);

/* --- PolyGas (src/PolyGas.hh:23-26), TTS (src/TTS.hh:23-25) ------------ */

/* TID_CALCSTATEHALF: EOS at the half step (reads ZR, ZVOLP, ZVOL0, ZE,
 * ZWRATE, ZM; writes ZP, ZSS).  Replaces gpu_state_half,
 * src/PolyGas.cu:27-101 (src/PolyGas.cc:63-108). */
int pennant_hip_calc_state_half(const pennant_mesh* m, const double* zr,
                                const double* zvolp, const double* zvol0,
                                const double* ze, const double* zwrate, const double* zm,
                                double* zp, double* zss, double gamma, double ssmin,
                                const double* dt, hipStream_t stream);

/* TID_CALCFORCEPGAS: sf = -zp[z] * ssurf.
 * Replaces gpu_calc_force_pgas, src/PolyGas.cu:103-142 (src/PolyGas.cc:164-185). */
int pennant_hip_calc_force_pgas(const pennant_mesh* m, pennant_cvec ssurf,
                                const double* zp, pennant_vec sf, hipStream_t stream);

/* TID_CALCFORCETTS: sf = -alfa*max(zss,ssmin)^2*(zr*smf*zarea/sarea - zr)*ssurf.
 * Replaces gpu_calc_force_tts, src/TTS.cu:22-95 (src/TTS.cc:55-104). */
int pennant_hip_calc_force_tts(const pennant_mesh* m, const double* sarea,
                               const double* smf, pennant_cvec ssurf, const double* zarea,
                               const double* zr, const double* zss, pennant_vec sf,
                               double alfa, double ssmin, hipStream_t stream);

/* --- QCS (src/QCS.hh:25-30) --------------------------------------------- */

/* TID_SETCORNERDIV: zuc, carea, ccos, cdiv, cevol, cdu.
 * Replaces gpu_zone_centered_velocity + gpu_corner_divergence,
 * src/QCS.cu:42-228 (src/QCS.cc:116-242). */
int pennant_hip_set_corner_div(const pennant_mesh* m, pennant_cvec ex, const double* elen,
                               pennant_cvec zx, pennant_cvec px, pennant_cvec pu,
                               pennant_vec zuc, double* carea, double* ccos, double* cdiv,
                               double* cevol, double* cdu, hipStream_t stream);

/* TID_SETQCNFORCE: crmu, cqe1, cqe2.
 * Replaces gpu_compute_crmu + gpu_compute_cqe, src/QCS.cu:230-355
 * (src/QCS.cc:246-327). */
int pennant_hip_set_qcn_force(const pennant_mesh* m, const double* elen, const double* cdiv,
                              const double* cdu, const double* cevol, const double* zrp,
                              const double* zss, pennant_cvec pu, double* crmu,
                              pennant_vec cqe1, pennant_vec cqe2, double qgamma,
                              double q1, double q2, hipStream_t stream);

/* TID_SETFORCEQCS: cw, clamped ccos (read-write), sfq.
 * Replaces gpu_set_force1 + gpu_set_force2, src/QCS.cu:357-440
 * (src/QCS.cc:331-380). */
int pennant_hip_set_force_qcs(const pennant_mesh* m, const double* carea,
                              pennant_cvec cqe1, pennant_cvec cqe2, const double* elen,
                              double* ccos, double* cw, pennant_vec sfq, hipStream_t stream);

/* TID_SETVELDIFF: ztmp = max(0, max_s |dot(du,dx)|/elen); zdu = q1*zss + 2*q2*ztmp.
 * Replaces gpu_set_vel_diff1 + gpu_set_vel_diff2, src/QCS.cu:442-549
 * (src/QCS.cc:384-448). */
int pennant_hip_set_vel_diff(const pennant_mesh* m, const double* elen, const double* zss,
                             pennant_cvec px, pennant_cvec pu, double* ztmp, double* zdu,
                             double q1, double q2, hipStream_t stream);

/* --- HydroBC (src/HydroBC.hh:30-35) ------------------------------------- */

/* TID_APPLYFIXEDBC: for the numb listed points, pu0 and pf are projected onto
 * the plane with unit normal vfix: v -= dot(v, vfix)*vfix.
 * Replaces gpu_apply_fixed_bc, src/HydroBC.cu:22-80 (src/HydroBC.cc:150-183). */
int pennant_hip_apply_fixed_bc(const int32_t* mapbp, int32_t numb, double vfix_x,
                               double vfix_y, pennant_vec pf, pennant_vec pu0,
                               hipStream_t stream);

/* --- Driver (src/Driver.hh:21-25) --------------------------------------- */

/* TID_CALCGLOBALDT on the device (src/Driver.cc:213-245):
 * dt = dtmax; cycle 0: min(dt, dtinit), else min(dt, dtfac * *dtlast);
 * min(dt, tstop - *time); cycle > 0: min(dt, *dtrec). */
int pennant_hip_calc_global_dt(const double* dtlast, const double* dtrec,
                               const double* time, int32_t cycle, double dtmax,
                               double dtinit, double dtfac, double tstop, double* dt,
                               hipStream_t stream);

/* ======================================================================== */
/* Driver API (whole problem on one device).                                */

typedef struct pennant_problem pennant_problem;

enum pennant_mode {
    PENNANT_MODE_TASKS = 0,   /* one launch per task ID, src/Hydro.cc:342-983 order */
    PENNANT_MODE_FUSED = 1,   /* fused cycle kernels (same arithmetic, fewer HBM passes) */
    PENNANT_MODE_MIXED = 2,   /* fused cycle, mixed precision: FP32 side/corner/QCS math,
                                 FP64 positions, volumes, EOS state, energy, work and dt */
    PENNANT_MODE_LEARNED = 3  /* the mixed form with the learned QCS coefficient
                                 (pennant_set_closure; exact QCS outside its envelope) */
};

/* Parses a .pnt deck (src/InputFile.cc:24-104 syntax), generates the mesh
 * (rect or pie, one piece), initialises it (src/Mesh.cc:616-623,
 * src/Hydro.cc:241-339) and uploads everything to the current device.
 * Returns NULL on error with a message in errbuf (the reference's fatal input
 * errors, src/GenMesh.cc:89-127, src/Driver.cc:80-83, src/Mesh.cc:220-223,
 * src/InputFile.cc:47-49, become this error return). */
pennant_problem* pennant_create(const char* deck_text, char* errbuf, int errlen);
void pennant_destroy(pennant_problem* pb);

/* Runs up to ncycles cycles (stopping at cstop or tstop) with the given mode,
 * asynchronously on the problem's stream.  Returns 0 or a hipError_t.
 *
 * With the tuning key "pipeline" set, the mixed modes run pipelined: one
 * zone pass holds a cycle's corrector and the next cycle's predictor, which
 * assumes the next dt computed without the volume-change limit; the
 * finalize checks that dt bit for bit and redoes the predictor when the
 * limit binds (every cycle while a shock crosses zones faster than cflv
 * allows, as in Sedov; seldom in smooth flow).  Results are bit-identical to
 * the unpipelined cycle.  The first pipelined launch after a change of state
 * (pennant_set_field, pennant_set_scalars, pennant_field_device, a closure,
 * capture or tuning change, cycles of another mode) computes the current
 * cycle's predictor first. */
int pennant_launch_cycles(pennant_problem* pb, int ncycles, int mode);

/* Number of look-ahead predictors the pipelined mixed cycle has redone
 * because the volume-change limit set the next dt.  Synchronous. */
int pennant_lookahead_redos(pennant_problem* pb, int64_t* count);

/* Waits for the stream; returns 0, PENNANT_EBADSIDES (message in
 * pennant_error) or a hipError_t. */
int pennant_sync(pennant_problem* pb);

/* Runs up to ncycles cycles and waits for them; returns the number of cycles
 * run or a negative error code (PENNANT_EBADSIDES, or a hipError_t negated).
 * Cycles are launched in chunks (32 growing to
 * 1024) with a stream sync after each, so a run that stops at cstop or tstop
 * long before ncycles does not queue the remaining (no-op) launches. */
int pennant_run(pennant_problem* pb, int ncycles, int mode);

const char* pennant_error(pennant_problem* pb);

/* Driver::run (src/Driver.cc:102-187): runs the problem in `mode` to its stop
 * condition (cstop, or tstop through the device predicate), reporting as the
 * reference does: after loop cycle 0 and after every cycle whose count is a
 * multiple of dtreport,
 *   "End cycle %6d, time = %11.5g, dt = %11.5g, wall = %11lld us\n"
 * (src/Driver.cc:147-156,330-344; wall = microseconds since the previous
 * report), then the summary of src/Driver.cc:169-180 ("Run complete", cycle
 * and cstop, time and tstop, "hydro cycle run time= %14.8g us" between
 * asterisk rules).  Each piece of text goes to report(text, user), or to
 * stdout when report is NULL.  The host reads the device control block only
 * at report boundaries (an event and an asynchronous copy each, up to 32
 * boundaries in flight), never per cycle.  *run_us (may be NULL) receives
 * the hydro cycle run time: HIP events from the call's start to the end of
 * the last cycle that ran.  Returns the number of cycles run, or a negative
 * error code: PENNANT_EBADSIDES (message in pennant_error), PENNANT_EINVAL,
 * or a hipError_t negated. */
typedef void (*pennant_report_fn)(const char* text, void* user);
int pennant_drive(pennant_problem* pb, int mode, pennant_report_fn report, void* user,
                  double* run_us);

/* Counts: out[0]=nump, out[1]=numz, out[2]=nums, out[3]=number of BC lists. */
void pennant_sizes(pennant_problem* pb, int64_t* out4);

/* Device control block, copied to the host (synchronises):
 * out[0]=time, out[1]=dt (of the last cycle run), out[2]=dtrec, out[3]=cycle,
 * out[4]=dtnew, out[5]=dvovmax, out[6]=numsbad_pred, out[7]=numsbad_corr,
 * out[8]=next dt (calcGlobalDt for the next cycle), out[9]=done flag. */
int pennant_scalars(pennant_problem* pb, double* out10);
int pennant_set_scalars(pennant_problem* pb, const double* in10);

/* Field access by reference field name ("px", "pu0", "zr", "sfq", ...;
 * lower-case FID names, plus "exp" for EXP).  ncomp is 1 or 2; data is
 * interleaved (x0,y0,x1,y1,...) for vector fields.  Synchronous.
 * The mixed modes do not store zr and ze each cycle (their predictors derive
 * them from zm, zvol and zetot); these calls, the writers, pennant_energy
 * and FP64 or task-mode cycles first store them, with the corrector's
 * quotients.  A written zr or ze is what the next predictor reads. */
int pennant_field_info(pennant_problem* pb, const char* name, int64_t* count, int* ncomp);
int pennant_get_field(pennant_problem* pb, const char* name, double* host_out);
int pennant_set_field(pennant_problem* pb, const char* name, const double* host_in);
/* device pointer of one component (0 = x or scalar, 1 = y), with zr and ze
 * stored (on the problem's stream); the caller may write through it, so the
 * pipelined cycle's look-ahead is recomputed.  NULL for an unknown name or a
 * launch error. */
double* pennant_field_device(pennant_problem* pb, const char* name, int comp);

/* Maps as int32 host copies: "mapsp1", "mapsp2", "mapsz", "mapss3", "mapss4",
 * "znump", "zsbeg", "pcbeg", "pcside". */
int pennant_get_map(pennant_problem* pb, const char* name, int32_t* host_out, int64_t* count);

/* The device mesh descriptor (device pointers) and the problem's stream, for
 * callers that drive the task entry points themselves. */
const pennant_mesh* pennant_mesh_desc(pennant_problem* pb);
hipStream_t pennant_stream(pennant_problem* pb);

/* One doCycle step in task mode (steps numbered as in src/Hydro.cc:342-983:
 * 0 copy zvol0, 1 point init + advPosHalf, 2 calcCtrs pred, 3 calcVols pred,
 * 4 calcSurfVecs, 5 calcEdgeLen, 6 calcCharLen, 7 calcRho pred,
 * 8 calcCrnrMass, 9 calcStateHalf, 10 calcForcePGas, 11 calcForceTTS,
 * 12 setCornerDiv, 13 setQCnForce, 14 setForceQCS, 15 setVelDiff,
 * 16 sumCrnrForce, 17 applyFixedBC (all lists), 18 checkBadSides pred,
 * 19 calcAccel + advPosFull, 20 calcCtrs corr, 21 calcVols corr, 22 calcWork,
 * 23 calcWorkRate, 24 calcEnergy, 25 calcRho corr, 26 calcDtNew, 27 calcDvol,
 * 28 calcDt, 29 checkBadSides corr), using the control block's dt.
 * begin_cycle runs calcGlobalDt into the control block; end_cycle advances
 * time and cycle.  Synchronous; for step-by-step parity checks. */
int pennant_begin_cycle(pennant_problem* pb);
int pennant_step(pennant_problem* pb, int step);
int pennant_end_cycle(pennant_problem* pb);

/* Runs ncycles cycles in `mode` between HIP events recorded on the problem's
 * stream, then synchronises.  *total_ms receives the elapsed time.  In fused
 * mode, phase_ms (5 entries, may be NULL) receives the summed durations of
 * the zone-predictor, point, zone-corrector and finalize kernels and of the
 * learned pass (0 outside the learned form), measured by events recorded
 * between their launches in every cycle (direct launches then, not the
 * captured graph).  For the pipelined mixed cycle the first entry is the
 * zone pass (corrector and next predictor), the third is 0 and the fourth
 * includes the predictor redo.
 * Returns 0, PENNANT_EBADSIDES or a hipError_t. */
int pennant_time_cycles(pennant_problem* pb, int ncycles, int mode, double* total_ms,
                        double* phase_ms);

/* Measurement helper: copies nbytes between two device buffers `reps` times
 * with a plain streaming kernel (k_stream_copy) and reports the average time
 * per copy.  Under a hardware-counter profiler its dispatches read and write
 * exactly nbytes each, which calibrates the FETCH_SIZE / WRITE_SIZE counters
 * against known traffic.  Allocates and frees its own buffers. */
int pennant_bench_copy(size_t nbytes, int reps, double* ms_per_copy);

/* Launch-shape choices of the fused cycle; results do not depend on them.
 *   "k1_layout"  0: one zone per thread (default), 1: one corner per lane
 *   "max_blocks" grid cap of the grid-stride zone kernels, 1..4096 (4096)
 *   "xcd"        1: each of the 8 XCDs takes a contiguous band of zones (default),
 *                0: plain grid-stride order
 *   "graph"      1: pennant_launch_cycles replays runs of 16 fused cycles as one
 *                captured HIP graph (default), 0: direct launches
 *   "pipeline"   1: the mixed modes run the pipelined cycle, 0: the zone
 *                predictor and corrector as separate passes (default);
 *                pennant_autotune tries both for the mixed modes
 *   "rect_maps"  1: on a rect mesh in the generator's numbering, the zone
 *                kernels compute each zone's points and corner slots from
 *                its index instead of reading mapsp1 and the slot map
 *                (default; pennant_get_tuning reports 1 only when the mesh
 *                qualifies), 0: they read the maps
 * pennant_set_tuning returns PENNANT_EINVAL for an unknown key or value. */
int pennant_set_tuning(pennant_problem* pb, const char* key, int value);
int pennant_get_tuning(pennant_problem* pb, const char* key);
/* Autotuner: times `cycles` cycles of `mode` at every launch shape (K1
 * layout for the FP64 form, grid cap 256..4096, XCD mapping) from one
 * snapshot of the state, keeps the fastest and restores the snapshot (the
 * problem does not advance).  best_ms (may be NULL) receives its time. */
int pennant_autotune(pennant_problem* pb, int mode, int cycles, double* best_ms);

/* Writes <path> in the reference's .xy format (src/WriteXY.cc:28-53). */
int pennant_write_xy(pennant_problem* pb, const char* path);

/* Writes the EnSight Gold set <base>.case, <base>.geo, <base>.zr, <base>.ze,
 * <base>.zp of the problem's current state (cycle, time, px, zr, ze, zp), in
 * the reference's format (src/ExportGold.cc:31-307, called from
 * Mesh::writeTask, src/Mesh.cc:1570-1604). */
int pennant_write_gold(pennant_problem* pb, const char* base);

/* The same writer on host arrays (no device needed): px/py are the point
 * coordinates, znump and mapsp1 the mesh maps (zone z's sides are contiguous,
 * starting at the sum of znump over the zones before z).  Returns 0, or
 * PENNANT_EINVAL when a file cannot be opened. */
int pennant_write_gold_host(const char* base, int32_t cycle, double time, const double* zr,
                            const double* ze, const double* zp, const int32_t* znump,
                            int32_t numz, const double* px, const double* py, int32_t nump,
                            const int32_t* mapsp1);

/* Host-only view of the deck and generated mesh (no device needed): the same
 * parsing, error messages and numbering pennant_create uses.  Names for
 * pennant_host_mesh_get: the int32 maps of pennant_get_map, and "px", "py"
 * (double point coordinates). */
typedef struct pennant_host_mesh pennant_host_mesh;
pennant_host_mesh* pennant_host_mesh_create(const char* deck_text, char* errbuf, int errlen);
int pennant_host_mesh_get(pennant_host_mesh* hm, const char* name, void* host_out, int64_t* count);
void pennant_host_mesh_destroy(pennant_host_mesh* hm);

/* ======================================================================== */
/* Learned QCS closure.  No counterpart in the reference: it replaces the    */
/* Kurapatenko coefficient of setQCnForce (src/QCS.cc:285-296) by a small    */
/* network, 10 per-corner features -> 64 -> 64 -> 2 coefficients, BF16      */
/* weights and activations with FP32 accumulation (MFMA).  The features,     */
/* outputs and training targets are defined in csrc/closure.h.               */

typedef struct pennant_mlp pennant_mlp;

/* Number of FP32 parameters: W1[64][16] (input columns 0-9 features, 10 the
 * bias, 11-15 zero), W2[64][64], b2[64], W3[2][64], b3[2]. */
int pennant_mlp_num_params(void);
/* He-uniform weights from `seed`, zero biases, an unbounded envelope. */
pennant_mlp* pennant_mlp_create(uint64_t seed);
void pennant_mlp_destroy(pennant_mlp* m);
int pennant_mlp_set_params(pennant_mlp* m, const float* host, int n);   /* resets Adam */
int pennant_mlp_get_params(pennant_mlp* m, float* host, int n);
/* Feature envelope: a zone whose compressing corners have a feature outside
 * [lo[k], hi[k]] uses the exact QCS coefficient instead of the network. */
int pennant_mlp_set_envelope(pennant_mlp* m, const float* lo10, const float* hi10);
int pennant_mlp_get_envelope(pennant_mlp* m, float* lo10, float* hi10);
/* Batched inference on device arrays x (n x 10) -> y (n x 2).  path 0: the
 * MFMA GEMM kernels used by training; path 1: the wave-level routine the
 * fused predictor kernel runs. */
int pennant_mlp_forward(pennant_mlp* m, const float* x, int64_t n, float* y, int path);
/* Loss (mean over the n tuples of the summed squared output error) and its
 * gradient with respect to every parameter, on device tuples (n x 12: 10
 * features, 2 targets); grad is a host array of pennant_mlp_num_params(). */
int pennant_mlp_loss_grad(pennant_mlp* m, const float* tuples, int64_t n, float* grad, float* loss);
/* `steps` Adam steps (lr, beta 0.9/0.999) on minibatches of `batch` tuples
 * drawn with replacement from the device tuples by a generator seeded with
 * `seed`; loss_out (host, may be NULL) receives each step's minibatch loss. */
int pennant_mlp_train(pennant_mlp* m, const float* tuples, int64_t n, int steps, int batch,
                      float lr, uint64_t seed, float* loss_out);

/* Attaches a trained network to a problem (NULL detaches); cycles in
 * PENNANT_MODE_LEARNED use it.  The problem keeps a pointer: the network
 * must outlive the problem's learned cycles. */
int pennant_set_closure(pennant_problem* pb, pennant_mlp* m);
/* Counters of learned cycles since the problem was created: out[0] zones
 * with compressing corners evaluated by the network, out[1] such zones that
 * fell back to the exact coefficient (a feature outside the envelope). */
int pennant_closure_stats(pennant_problem* pb, int64_t* out2);

/* Training data: while a capture buffer is set (device, cap x 12 floats),
 * every PENNANT_MODE_MIXED cycle appends one tuple (features, exact target
 * coefficients) per compressing corner at its predictor half step; tuples
 * past `cap` are counted but not stored.  The order of the tuples within a
 * cycle is not fixed.  pennant_set_capture(pb, NULL, 0) stops capturing. */
int pennant_set_capture(pennant_problem* pb, float* dev_buf, int64_t cap);
/* Tuples produced since the capture buffer was set (may exceed cap). */
int pennant_capture_count(pennant_problem* pb, int64_t* count);

/* Total energy of the current state (reported, not used by the cycle):
 * internal = sum_z zetot; kinetic = sum over corners of
 * zm * (zarea * x_p * (smf_s + smf_s3) / 2) / zvol * |u_p|^2 / 2, the
 * corner-volume-weighted kinetic energy of each zone (RZ volumes per radian,
 * as zm and zvol). */
int pennant_energy(pennant_problem* pb, double* internal, double* kinetic);

/* ======================================================================== */
/* Batched problems.  No counterpart in the reference (its closest analogue  */
/* is a Legion index launch over the pieces of one problem,                  */
/* src/Hydro.cc:385): many independent problems of uneven sizes advance      */
/* together, one launch per phase over all of them, each with its own dt,    */
/* parameters and stop condition.  Modes PENNANT_MODE_FUSED, _MIXED and      */
/* _LEARNED; zones must have 3 or 4 sides (rect and pie meshes).             */

typedef struct pennant_batch pennant_batch;

/* Creates and initialises every deck (as pennant_create does), then moves
 * them into one arena on the current device.  NULL on error (message in
 * errbuf, naming the deck). */
pennant_batch* pennant_batch_create(const char* const* decks, int n, char* errbuf, int errlen);
void pennant_batch_destroy(pennant_batch* b);
int pennant_batch_count(pennant_batch* b);
/* out3: nump, numz, nums of problem i. */
int pennant_batch_sizes(pennant_batch* b, int i, int64_t* out3);
const char* pennant_batch_error(pennant_batch* b);
/* Queues ncycles cycles (asynchronous; problems that have stopped skip
 * them).  Runs of 16 cycles replay a captured HIP graph unless the tuning
 * key "graph" is 0. */
int pennant_batch_launch_cycles(pennant_batch* b, int ncycles, int mode);
/* Waits; PENNANT_EBADSIDES (message in pennant_batch_error) when a problem
 * stopped on negative side volumes. */
int pennant_batch_sync(pennant_batch* b);
/* Waits; number of problems still running. */
int pennant_batch_active(pennant_batch* b);
int pennant_batch_time_cycles(pennant_batch* b, int ncycles, int mode, double* total_ms);
/* The scalars of problem i (the layout of pennant_scalars). */
int pennant_batch_scalars(pennant_batch* b, int i, double* out10);
/* Field of problem i: "px", "pu" (interleaved x, y), "zr", "ze", "zp",
 * "zetot", "zwrate", "zvol", "zm". */
int pennant_batch_get_field(pennant_batch* b, int i, const char* name, double* out);
int pennant_batch_set_closure(pennant_batch* b, pennant_mlp* m);
int pennant_batch_closure_stats(pennant_batch* b, int64_t* out2);
/* Training-tuple capture over all problems (see pennant_set_capture). */
int pennant_batch_set_capture(pennant_batch* b, float* dev_buf, int64_t cap);
int pennant_batch_capture_count(pennant_batch* b, int64_t* count);
/* "max_blocks" (1..65535, default 2048), "xcd" (0/1), "graph" (0/1),
 * "rect_maps" (0/1, default 1: the zone kernels compute the points and corner
 * slots of the batch's rect problems, as pennant_set_tuning's "rect_maps"). */
int pennant_batch_set_tuning(pennant_batch* b, const char* key, int value);
int pennant_batch_get_tuning(pennant_batch* b, const char* key);
/* Times `cycles` cycles at each grid shape (grid cap x XCD mapping) from the
 * same snapshot of the state, keeps the fastest and restores the snapshot:
 * the problems do not advance.  best_ms (may be NULL) receives its time. */
int pennant_batch_autotune(pennant_batch* b, int mode, int cycles, double* best_ms);
/* Parameters of an ensemble of `count` problems: for k even a Sedov blast on
 * an n[k] x n[k] rect mesh with energy param[k] in [0.1, 1.0], for k odd a
 * Noh implosion on an n[k] x n[k] pie mesh with inflow speed param[k] in
 * [0.5, 2.0]; n log-uniform in [32, 512]; std::mt19937_64(seed), drawn in
 * the order n, param for each k. */
int pennant_ensemble_params(uint64_t seed, int count, int32_t* n, double* param);

/* ======================================================================== */
/* One mesh as numpcs pieces on one device: the reference's -n <numpcs>     */
/* (src/main.cc:61-80).  Pieces are near-square zone blocks                 */
/* (src/Mesh.hh:760-791; zone and point ownership src/GenMesh.cc:790-975).  */
/* Each piece holds its zones, their sides and every point they reach,      */
/* numbered private points, master points (owned, reached by other pieces   */
/* too), ghost points (owned elsewhere) -- the partition of                 */
/* src/Mesh.cc:432-617.  A cycle runs the fused kernels per piece; corners  */
/* at ghost points go to their owner, which sums a master point's corners   */
/* from every piece in ascending one-piece side order (the reference's      */
/* shared-point reductions, src/Hydro.cc:571-574,766-769, whose fold order  */
/* Legion leaves open), and the owners' new point state goes back to the    */
/* ghosts (the reads through the shared partition, src/Hydro.cc:450-455,    */
/* 866-867,902-906).  Results are bit-identical to the one-piece run.       */

typedef struct pennant_pieces pennant_pieces;

/* Parses the deck, partitions the mesh into numpcs pieces (pieces the block
 * rounding leaves empty are dropped) and initialises every piece.  NULL on
 * error with a message in errbuf. */
pennant_pieces* pennant_pieces_create(const char* deck_text, int numpcs, char* errbuf, int errlen);
/* The partition alone, computed on the host (no device needed): out[0]
 * numpcx, out[1] numpcy, out[2] the number n of non-empty pieces, then 7
 * values per piece: piece number, zones, sides, private, master and ghost
 * points, corners received.  Returns 3 + 7 n, or PENNANT_EINVAL (message in
 * errbuf) for an invalid deck or an out array shorter than outlen needs. */
int pennant_pieces_plan(const char* deck_text, int numpcs, int64_t* out, int64_t outlen, char* errbuf,
                        int errlen);
void pennant_pieces_destroy(pennant_pieces* pp);
const char* pennant_pieces_error(pennant_pieces* pp);
/* Number of (non-empty) pieces. */
int pennant_pieces_count(pennant_pieces* pp);
/* Piece i: out[0] piece number (py * numpcx + px), out[1] zones, out[2]
 * sides, out[3] private points, out[4] master points, out[5] ghost points,
 * out[6] corners its master points receive from other pieces, out[7]
 * numpcx * 65536 + numpcy. */
int pennant_pieces_layout(pennant_pieces* pp, int piece, int64_t* out8);
/* Runs up to ncycles cycles in PENNANT_MODE_FUSED or PENNANT_MODE_MIXED;
 * returns the cycles run, PENNANT_EBADSIDES or a negative HIP error. */
int pennant_pieces_run(pennant_pieces* pp, int ncycles, int mode);
/* Queues ncycles cycles between two events on the pieces' stream and waits;
 * *total_ms receives their time. */
int pennant_pieces_time_cycles(pennant_pieces* pp, int ncycles, int mode, double* total_ms);
/* The shared control block (layout of pennant_scalars). */
int pennant_pieces_scalars(pennant_pieces* pp, double* out10);
/* Entries (whole mesh) and components of a field, as pennant_field_info. */
int pennant_pieces_field_info(pennant_pieces* pp, const char* name, int64_t* count, int* ncomp);
/* A zone, side or point field in the one-piece numbering (points from their
 * owner piece; vector fields interleaved x, y). */
int pennant_pieces_get_field(pennant_pieces* pp, const char* name, double* out);

#ifdef __cplusplus
}
#endif

#endif /* PENNANT_HIP_H_ */
