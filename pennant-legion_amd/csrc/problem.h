// Host-side problem state shared by the driver translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "pennant_hip.h"

namespace pn {

// Device control block: the per-cycle scalars the reference threads through
// futures (src/Driver.cc:127-158, src/Hydro.cc:952-977).  `dt` is the dt of
// the cycle about to run; kernels read it from here so that no host sync is
// needed between cycles.
struct Ctl {
    double time;
    double dt;        // dt for the next cycle (calcGlobalDt result)
    double dtlast;    // dt of the last completed cycle
    double dtrec;     // calcDt result of the last completed cycle
    double dtnew;
    double dvovmax;
    int cycle;
    int done;         // 1 once cycle >= cstop, time >= tstop, or bad sides
    int numsbad_pred;
    int numsbad_corr;
    int err_cycle;    // cycle whose bad-side check failed, or -1
    int err_pred;     // 1 when the failed check was the predictor's
    // look-ahead of the pipelined mixed cycle: the predictor of the next
    // cycle (its dt limit and bad-side count), and whether it must be redone
    double la_dtnew;
    int la_bp;
    int redo;
    int la_redos;     // redone look-ahead predictors since the problem was created
    // 1: zr and ze in memory are current (after initialisation, FP64 cycles
    // and field writes) and the next predictor reads them; 0: the mixed
    // corrector stored neither, and predictors derive them from zm, zvol and
    // zetot with the same quotients until they are materialised
    int zre_valid;
};

#define PN_MAX_GRID 4096   // largest grid of a fused K1/K3 launch

// Slot of a corner record (K1 writes it, K2 reads it): position c of point
// p's corner list (ascending side order; k = pcbeg[p] + c is its point-corner
// CSR index) in a mesh of nump points.  The first four positions are
// corner-major, c * nump + p: neighbouring zones write the records of
// neighbouring points at one position, and neighbouring points read them,
// so both sides touch whole cache lines.  Later positions (points with more
// than four corners) take 4 * nump + k.  A mesh needs corner_slots() slots.
inline int64_t corner_slot(int64_t nump, int64_t p, int64_t c, int64_t k) {
    return c < 4 ? c * nump + p : 4 * nump + k;
}
inline int64_t corner_slots(int64_t nump, int64_t nums) { return 4 * nump + nums; }
// The kernels index FP32 corner records {fx, fy, mass} as 3 * slot + j in int.
inline bool corner_slots_fit(int64_t nump, int64_t nums) {
    return 3 * corner_slots(nump, nums) < (int64_t)0x7fffffff;
}
#define PN_HEAVY 64        // points with more corners get a whole block in K2

// Launch-shape choices of the fused cycle (pennant_set_tuning).
struct Tuning {
    int k1_layout = 0;       // 0 zone per thread, 1 corner per lane
    int max_blocks = 4096;   // grid cap of the grid-stride K1/K3 launches
    int xcd = 1;             // XCD-aware unit mapping of the grid-stride launches
    int graph = 1;           // batched cycles through a captured HIP graph
    int pipeline = 0;        // mixed forms: next cycle's predictor in the corrector pass
    int rect_maps = 1;       // structured rect meshes: zone points and corner slots computed, not loaded
};

struct Params {
    int cstop; double tstop, dtmax, dtinit, dtfac; int dtreport;
    double cfl, cflv, rinit, einit, rinitsub, einitsub, uinitradial;
    std::vector<double> bcx, bcy, subregion;
    double gamma, ssmin, alfa, qgamma, q1, q2;
    std::string meshtype; long nzx, nzy; double lenx, leny;
};

struct DevField {
    int kind;        // 0 point, 1 zone, 2 side
    int ncomp;       // 1 or 2
    double* c[2];
};

struct HostMesh {
    int nump = 0, numz = 0, nums = 0, maxznump = 0;
    std::vector<int> mapsp1, mapsp2, mapsz, mapss3, mapss4, znump, zsbeg, pcbeg, pcside;
    std::vector<double> px, py;
    bool cyclic = true;   // sides of each zone in cyclic order (fused path requirement)
};

struct BCList {
    double vx, vy;
    int numb;
    int* d_pts;
};

}  // namespace pn

struct pennant_problem {
    pn::Params prm;
    pn::HostMesh hm;
    pennant_mesh dm;                 // device pointers
    std::vector<void*> allocs;
    std::map<std::string, pn::DevField> fields;
    std::vector<pn::BCList> bcs;
    pn::Ctl* ctl = nullptr;          // device
    pn::Ctl* ctl_host = nullptr;     // pinned mirror
    void* ws = nullptr;              // reduction workspace
    hipStream_t stream = nullptr;
    std::string err;

    // fused-path buffers
    double* cfx = nullptr; double* cfy = nullptr; double* cmass = nullptr;
    double* sfwx = nullptr; double* sfwy = nullptr;
    double* vs2 = nullptr; double* pxpx = nullptr;   // vs2: {x, y} per point
    float* fcrn = nullptr;           // mixed form: per corner slot {fx, fy, mass}
    float* fsfw = nullptr;           // mixed form: per side {x, y}
    const float* fsmf = nullptr;
    unsigned char* pbc = nullptr;    // per-point BC membership bits (one per list)
    double* part_d = nullptr;        // block partials: dt minima, then dvol maxima
    int* part_i = nullptr;           // block partials: bad-side counts (pred, corr)
    int npart_pred = 0, npart_corr = 0;
    int* pcpos = nullptr;            // side -> slot of its corner record (corner_slot)
    bool smf_stale = false;          // smf may have been written through its device pointer
    int* heavy = nullptr;            // points with more corners than a K2 thread takes
    int nheavy = 0;
    // learned closure (a ClosurePack on the device) and training-tuple capture
    const void* closure = nullptr;
    float* cap_buf = nullptr;
    int64_t cap = 0;
    // [0] tuples, [1] learned zones, [2] fallback zones; [3], [4] the same
    // counts of a look-ahead predictor, added to [1], [2] when its cycle ends
    unsigned long long* cl_counters = nullptr;
    // structured rect mesh in the generator's numbering (rect_detect): nzx,
    // else 0, and the multiplier and shift of z / nzx
    int rect_nzx = 0;
    unsigned long long rect_m = 0;
    int rect_sh = 0;
    int* zlist = nullptr;            // zones for the learned pass (per run, at the run's zone range)
    unsigned* zcount = nullptr;      // per run and list slice, one 128-byte line each
    pn::Tuning tune;
    hipGraphExec_t gexec = nullptr;  // captured fused cycles of form graph_form
    int graph_form = -1;             // form, + 16 for the pipelined cycle
    // pipelined mixed cycle: the form whose look-ahead predictor results
    // (corner records, side work forces, zpn, Ctl la_*) hold for the current
    // cycle, or -1 when they must be computed first
    int la_form = -1;
    double* zpn = nullptr;           // the look-ahead predictor's zp
    bool zre_stale = false;          // mixed cycles ran since zr, ze were last materialised
    std::vector<int> zruns;          // [zbeg, zend, nsides] triples with uniform nsides
    // a piece of a multi-piece run: the point pass updates points [0, npown)
    // (private and master points); -1: every point
    int npown = -1;
    int nbcs = 0;
    double bcvx[8], bcvy[8];
    double* bcv_dev = nullptr;       // device copy: bcvx then bcvy

    double* f(const char* name, int comp = 0) { return fields.at(name).c[comp]; }
    pennant_cvec cv(const char* name) { pn::DevField& d = fields.at(name); pennant_cvec v = {d.c[0], d.c[1], 1}; return v; }
    pennant_vec vv(const char* name) { pn::DevField& d = fields.at(name); pennant_vec v = {d.c[0], d.c[1], 1}; return v; }
};

namespace pn {
// Batched problems: zone units (at most PN_BLOCK zones of one side-count run
// of one problem), point units (at most PN_BLOCK points of one problem),
// heavy points and the per-problem parameters.
struct BUnit {
    int zb, ze, sb, ns, pid, pad0, pad1, pad2;
};
struct PUnit {
    int pb, pe, pid, pad;
};
struct BHeavy {
    int p, pid;
};
struct BProb {
    double gamma, ssmin, alfa, qgamma, q1, q2, cfl, cflv, dtmax, dtfac, tstop;
    double bcvx[8], bcvy[8];
    int cstop, nbcs, ubeg, uend;
    // rect problems (rect_nzx > 0): z / nzx = (z * rect_m) >> rect_sh for the
    // problem's own zone z; so, po: its first side and point in the arena
    unsigned long long rect_m;
    int rect_nzx, rect_sh, so, po;
};
}  // namespace pn

struct pennant_batch {
    int nprob = 0;
    int64_t numz = 0, nums = 0, nump = 0;
    std::vector<int64_t> zoff, soff, poff;   // per problem, plus the totals
    std::vector<void*> allocs;
    // state (arena: all problems concatenated; maps hold arena indices)
    double *pxx = nullptr, *pxy = nullptr, *pux = nullptr, *puy = nullptr;
    double *zr = nullptr, *ze = nullptr, *zetot = nullptr, *zwrate = nullptr, *zvol = nullptr;
    double *zp = nullptr, *zm = nullptr, *smf = nullptr;
    float* fsmf = nullptr;
    int *mapsp1 = nullptr, *pcbeg = nullptr, *pcpos = nullptr;
    unsigned char* pbc = nullptr;
    // fused intermediates
    double *cfx = nullptr, *cfy = nullptr, *cmass = nullptr, *sfwx = nullptr, *sfwy = nullptr;
    double *vs2 = nullptr, *pxpx = nullptr;
    float *fcrn = nullptr, *fsfw = nullptr;
    // work tables
    pn::BUnit* units = nullptr;
    int nunits = 0;
    pn::PUnit* punits = nullptr;
    int npunits = 0;
    pn::BHeavy* heavy = nullptr;
    int nheavy = 0;
    pn::BProb* probs = nullptr;
    pn::Ctl* ctls = nullptr;
    pn::Ctl* ctls_host = nullptr;                // pinned mirror
    double* part_d = nullptr;                    // per unit: dt minima, then dvol maxima
    int* part_i = nullptr;                       // per unit: bad sides (pred, corr)
    // closure and capture
    const void* closure = nullptr;
    float* cap_buf = nullptr;
    int64_t cap = 0;
    unsigned long long* cl_counters = nullptr;
    pn::Tuning tune;
    hipStream_t stream = nullptr;
    // captured cycles: graph_cycles cycles of form graph_form
    hipGraphExec_t gexec = nullptr;
    int graph_form = -1, graph_cycles = 0;
    bool zre_stale = false;          // mixed cycles ran since zr, ze were last materialised
    std::string err;
};

namespace pn {
// one cycle of every unfinished problem of a batch (fused.hip)
int batch_cycle(pennant_batch* b, int form);
int launch_to_float(const double* src, float* dst, int64_t n, hipStream_t s);
// fused cycle (fused.hip)
int fused_setup(pennant_problem* pb);
// forms of the fused cycle
enum { FORM_FP64 = 0, FORM_MIXED = 1, FORM_LEARNED = 2 };
// ev: NULL, or 6 events recorded before K1, the learned pass, K2, K3, K4 and
// after K4
int fused_cycle(pennant_problem* pb, hipEvent_t* ev, int form);
// Pipelined mixed cycle (FORM_MIXED, FORM_LEARNED): K2, then one zone pass
// with this cycle's corrector and the next cycle's predictor, K4, and a redo
// of the predictor when the next dt differs from the one it assumed.
// fused_prologue computes the current cycle's predictor first.
// ev: NULL, or 5 events recorded before K2, the zone pass, K4 and the redo,
// the learned pass, and after it
int fused_prologue(pennant_problem* pb, int form);
int fused_cycle_pipe(pennant_problem* pb, hipEvent_t* ev, int form);
// Stores zr = zm / zvol and ze = zetot / (zm + 1e-99) where mixed cycles left
// them to be derived (the mixed corrector does not store them).
int materialize_zre(pennant_problem* pb);
// Empties the learned pass's zone lists (a pipelined launch ends with the
// look-ahead predictor's list in them).
int reset_zone_lists(pennant_problem* pb);
// Rebuilds the FP32 copy of smf that the mixed forms read (at set-up and
// whenever smf may have changed).
int refresh_side_fracs(pennant_problem* pb);
int batch_materialize_zre(pennant_batch* b);
// device init and end-of-cycle kernels (driver_kernels.hip)
int launch_init_state(pennant_problem* pb);
int launch_end_cycle(pennant_problem* pb, int compute_dtrec);
int launch_first_dt(pennant_problem* pb);
// doCycle's copies and fills (src/Hydro.cc:361-368,389-418) as kernels, so
// that a captured task-mode cycle holds kernel nodes only: zvol0 = zvol;
// px0 = px, pu0 = pu, pmaswt = 0, pf = 0
int launch_copy_zvol0(pennant_problem* pb);
int launch_point_begin(pennant_problem* pb);
// Multi-piece runs (pieces.cpp, fused.hip): per-piece device pointers the
// exchange kernels use, and one exchange entry (element src of piece sp to
// element dst of piece dp).
struct PieceTab {
    double *pxx, *pxy, *pux, *puy, *vs2, *pxpx, *cfx, *cfy, *cmass;
    float* fcrn;
};
struct XEnt {
    int sp, src, dp, dst;
};
// deck parsing and the one-piece mesh generator (driver.cpp)
bool read_params(const char* text, Params& q, std::string& err);
void build_mesh(const Params& q, HostMesh& h);
// a problem on a prebuilt mesh (driver.cpp); npown as pennant_problem::npown
pennant_problem* create_problem(const Params& q, HostMesh&& h, int npown, std::string& err);
// points the point pass updates
inline int own_points(const pennant_problem* pb) {
    return pb->npown < 0 ? pb->hm.nump : pb->npown;
}
}  // namespace pn

// A mesh run as several pieces on one device (pennant_pieces_create).  Each
// piece is a problem of its own over its zones, their sides and every point
// they reach, numbered private points, master points (owned here, reached by
// other pieces too), then ghost points (owned by another piece).  All pieces
// share piece 0's control block and one array of block partials.
struct pennant_pieces {
    pn::Params prm;
    int numpcs = 0, numpcx = 0, numpcy = 0;
    std::vector<pennant_problem*> pc;             // the non-empty pieces, in piece order
    std::vector<int> pid;                         // each one's piece number, py * numpcx + px
    std::vector<std::vector<int>> zglob, sglob, pglob;   // local -> one-piece numbering
    std::vector<int> nprv, nmstr, nghost, nrecv;  // nrecv: corners received for master points
    int64_t numz = 0, nums = 0, nump = 0;         // the whole mesh
    pn::PieceTab* tab = nullptr;                  // device, one per piece
    pn::XEnt* halo = nullptr;                     // owner point -> ghost point
    int nhalo = 0;
    pn::XEnt* cx = nullptr;                       // corner slot at a ghost point -> owner's slot
    int ncx = 0;
    double* part_d = nullptr;                     // partials: [0, npart) K1's, [npart, 2 npart) K3's
    int* part_i = nullptr;
    int npart = 0;
    bool zre_stale = false;                       // mixed cycles ran since zr, ze were stored
    // the pieces' kernels of a phase run on their own streams (fork and join
    // events on piece 0's stream); runs of 16 cycles replay a captured graph
    std::vector<hipEvent_t> ev_fork, ev_join;     // 3 forks; 3 joins per piece
    hipGraphExec_t gexec = nullptr;
    int graph_form = -1;
    std::vector<void*> allocs;
    std::string err;
};

namespace pn {
// one cycle of every piece (fused.hip): K1 per piece, the corner exchange, K2
// per piece, the halo, K3 per piece, one finalize
int pieces_cycle(pennant_pieces* pp, int form);
// zr and ze stored in every piece where mixed cycles left them derived
int pieces_materialize_zre(pennant_pieces* pp);
}  // namespace pn
