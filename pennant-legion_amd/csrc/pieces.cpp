// A mesh run as several pieces on one device.
//
// The reference splits a run into numpcs pieces (-n <numpcs>, src/main.cc:61-80):
// near-square blocks of zones (src/Mesh.hh:760-791, src/GenMesh.cc:898-975),
// each point owned by one piece (src/GenMesh.cc:790-873), and the points
// partitioned into private points (reached by their owner only), master
// points (owned here and reached by other pieces) and, per piece, ghost
// points (reached here, owned elsewhere) (src/Mesh.cc:432-617).  Sides read
// ghost points through the shared partition (src/Hydro.cc:450-455,866-867,
// 902-906) and reduce their corner mass and force into shared points
// (src/Hydro.cc:571-574,766-769).
//
// Here every piece is a problem of its own, with its own arrays, over its
// zones (ascending one-piece order), their sides and every point they reach,
// numbered [private | master | ghost].  The corner list of a master point
// holds the corners of every piece that reaches it, in ascending one-piece
// side order; the corners at ghost points are sent to the owner's list by the
// corner exchange, and the owners' new point state goes back to the ghosts by
// the halo (pieces_cycle in fused.hip).  The reference leaves the fold order of
// shared-point reductions to Legion; here it is the one-piece order, so a
// run in pieces is bit-identical to the one-piece run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "problem.h"

using namespace pn;

namespace {

// Piece grid numpcx x numpcy with numpcx * numpcy = numpcs and blocks as
// close to square as the divisors of numpcs allow (the rule of
// src/Mesh.hh:760-791): the ideal count along the shorter side is
// sqrt(numpcs * short / long); the nearest divisors below and above it are
// tried, and the one whose blocks have the shorter long side wins (the lower
// one on a tie).
void piece_grid(long numpcs, long nzx, long nzy, long& numpcx, long& numpcy) {
    double a = (double)nzx, b = (double)nzy;
    const bool swapped = a > b;
    if (swapped) std::swap(a, b);
    const double ideal = std::sqrt((double)numpcs * a / b);
    long lo = std::max(1L, (long)std::floor(ideal + 1.e-12));
    while (numpcs % lo != 0) --lo;
    long hi = (long)std::ceil(ideal - 1.e-12);
    while (numpcs % hi != 0) ++hi;
    const double side_lo = std::max(a / (double)lo, b / (double)(numpcs / lo));
    const double side_hi = std::max(a / (double)hi, b / (double)(numpcs / hi));
    numpcx = (side_lo <= side_hi) ? lo : hi;
    numpcy = numpcs / numpcx;
    if (swapped) std::swap(numpcx, numpcy);
}

template <class T> T* upload_to(pennant_pieces* pp, const std::vector<T>& v) {
    void* p = nullptr;
    const size_t bytes = v.empty() ? sizeof(T) : v.size() * sizeof(T);
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    pp->allocs.push_back(p);
    if (!v.empty() && hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return nullptr;
    return static_cast<T*>(p);
}

int sync_ctl(pennant_pieces* pp) {
    pennant_problem* p0 = pp->pc[0];
    int rc = (int)hipMemcpyAsync(p0->ctl_host, p0->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, p0->stream);
    if (!rc) rc = (int)hipStreamSynchronize(p0->stream);
    return rc;
}

#define PN_PIECES_GRAPH_CYCLES 16

// Captures PN_PIECES_GRAPH_CYCLES cycles of `form` as one HIP graph (the
// pieces' kernels of a phase on parallel branches, kernel nodes only).
int build_pieces_graph(pennant_pieces* pp, int form) {
    if (pp->gexec) (void)hipGraphExecDestroy(pp->gexec);
    pp->gexec = nullptr;
    pp->graph_form = -1;
    hipStream_t s0 = pp->pc[0]->stream;
    hipGraph_t g = nullptr;
    int rc = (int)hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal);
    if (rc) return rc;
    for (int k = 0; k < PN_PIECES_GRAPH_CYCLES && !rc; ++k) rc = pieces_cycle(pp, form);
    const int rc2 = (int)hipStreamEndCapture(s0, &g);
    if (rc || rc2) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : rc2;
    }
    rc = (int)hipGraphInstantiate(&pp->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (rc) { pp->gexec = nullptr; return rc; }
    pp->graph_form = form;
    return 0;
}

// Queues ncycles cycles of `form`: runs of PN_PIECES_GRAPH_CYCLES replay the
// captured graph, the rest launch directly.
int launch_pieces_cycles(pennant_pieces* pp, int ncycles, int form) {
    int rc = 0;
    if (form == FORM_FP64 && (rc = pieces_materialize_zre(pp))) return rc;
    int left = ncycles;
    if (left >= PN_PIECES_GRAPH_CYCLES) {
        if (pp->graph_form != form || !pp->gexec) {
            if ((rc = build_pieces_graph(pp, form))) return rc;
        }
        for (; left >= PN_PIECES_GRAPH_CYCLES && !rc; left -= PN_PIECES_GRAPH_CYCLES)
            rc = (int)hipGraphLaunch(pp->gexec, pp->pc[0]->stream);
        if (form != FORM_FP64) pp->zre_stale = true;
    }
    for (; left > 0 && !rc; --left) rc = pieces_cycle(pp, form);
    return rc;
}

int fused_form_of(int mode) {
    switch (mode) {
        case PENNANT_MODE_FUSED: return FORM_FP64;
        case PENNANT_MODE_MIXED: return FORM_MIXED;
        default: return -1;
    }
}

// The partition of a mesh into pieces, computed on the host: which zones,
// sides and points each piece holds, their local numbering and meshes, and
// the corners that master points receive.
struct Recv {
    int owner, lp, c, k, gside;   // owner piece index, master point, position, CSR index, side
};
struct Plan {
    int numpcx = 1, numpcy = 1;
    std::vector<int> pid, own, idx_of_piece, slocal, spiece;
    std::vector<std::vector<int>> zglob, sglob, pglob, plocal;
    std::vector<int> nprv, nmstr, nghost, nrecv;
    std::vector<Recv> recvs;
    std::vector<HostMesh> meshes;
};

bool plan_pieces(const Params& q, const HostMesh& g, int numpcs, Plan& pl, std::string& err) {
    const bool pie = (q.meshtype == "pie");
    if (!pie && q.meshtype != "rect") { err = "meshtype " + q.meshtype + " is not supported"; return false; }
    const long nzx = q.nzx, nzy = q.nzy, npx = nzx + 1;
    long numpcx = 1, numpcy = 1;
    piece_grid(numpcs, nzx, nzy, numpcx, numpcy);
    pl.numpcx = (int)numpcx; pl.numpcy = (int)numpcy;
    const long zpx = (nzx + numpcx - 1) / numpcx, zpy = (nzy + numpcy - 1) / numpcy;

    // zone -> piece: the block holding its (i, j) (src/GenMesh.cc:898-975)
    std::vector<int> zpc((size_t)g.numz);
    for (long z = 0; z < g.numz; ++z) zpc[(size_t)z] = (int)((z / nzx) / zpy * numpcx + (z % nzx) / zpx);
    // point -> owner: the block of the zone whose first corner it is, the
    // last row and column folded into the blocks before them; the pie
    // origin belongs to piece 0 (src/GenMesh.cc:790-873)
    std::vector<int>& own = pl.own;
    own.assign((size_t)g.nump, 0);
    for (long p = 0; p < g.nump; ++p) {
        long i, j;
        if (!pie) { i = p % npx; j = p / npx; }
        else if (p == 0) { own[0] = 0; continue; }
        else { i = (p - 1) % npx; j = (p - 1) / npx + 1; }
        own[(size_t)p] = (int)(std::min(j, nzy - 1) / zpy * numpcx + std::min(i, nzx - 1) / zpx);
    }
    // shared points: reached by a piece other than their owner
    std::vector<char> shared((size_t)g.nump, 0), owner_reaches((size_t)g.nump, 0);
    for (int s = 0; s < g.nums; ++s) {
        const int p = g.mapsp1[(size_t)s], pc = zpc[(size_t)g.mapsz[(size_t)s]];
        if (pc != own[(size_t)p]) shared[(size_t)p] = 1;
        else owner_reaches[(size_t)p] = 1;
    }
    for (long p = 0; p < g.nump; ++p)
        if (!owner_reaches[(size_t)p]) { err = "a point is not reached by its owner piece"; return false; }

    // per piece: zones, sides and points in local order
    const int npc_all = (int)(numpcx * numpcy);
    std::vector<std::vector<int>> zl((size_t)npc_all);
    for (int z = 0; z < g.numz; ++z) zl[(size_t)zpc[(size_t)z]].push_back(z);
    pl.slocal.assign((size_t)g.nums, -1);      // global side -> local side in its piece
    pl.spiece.assign((size_t)g.nums, -1);      // global side -> piece index
    pl.idx_of_piece.assign((size_t)npc_all, -1);
    for (int c = 0; c < npc_all; ++c) {
        if (zl[(size_t)c].empty()) continue;            // pieces left empty by the block rounding
        pl.idx_of_piece[(size_t)c] = (int)pl.pid.size();
        pl.pid.push_back(c);
    }
    const int npc = (int)pl.pid.size();
    pl.zglob.resize((size_t)npc); pl.sglob.resize((size_t)npc); pl.pglob.resize((size_t)npc);
    pl.plocal.resize((size_t)npc);
    pl.nprv.assign((size_t)npc, 0); pl.nmstr.assign((size_t)npc, 0);
    pl.nghost.assign((size_t)npc, 0); pl.nrecv.assign((size_t)npc, 0);
    pl.meshes.resize((size_t)npc);
    for (int qi = 0; qi < npc; ++qi) {
        const int c = pl.pid[(size_t)qi];
        std::vector<int>& zg = pl.zglob[(size_t)qi];
        std::vector<int>& sg = pl.sglob[(size_t)qi];
        std::vector<int>& pg = pl.pglob[(size_t)qi];
        zg = zl[(size_t)c];
        for (int z : zg)
            for (int s = g.zsbeg[(size_t)z]; s < g.zsbeg[(size_t)z + 1]; ++s) {
                pl.slocal[(size_t)s] = (int)sg.size();
                pl.spiece[(size_t)s] = qi;
                sg.push_back(s);
            }
        std::vector<char> reach((size_t)g.nump, 0);
        for (int s : sg) reach[(size_t)g.mapsp1[(size_t)s]] = 1;
        std::vector<int> prv, mst, gst;
        for (int p = 0; p < g.nump; ++p) {
            if (!reach[(size_t)p]) continue;
            if (own[(size_t)p] != c) gst.push_back(p);
            else if (shared[(size_t)p]) mst.push_back(p);
            else prv.push_back(p);
        }
        pl.nprv[(size_t)qi] = (int)prv.size();
        pl.nmstr[(size_t)qi] = (int)mst.size();
        pl.nghost[(size_t)qi] = (int)gst.size();
        pg = prv;
        pg.insert(pg.end(), mst.begin(), mst.end());
        pg.insert(pg.end(), gst.begin(), gst.end());
        std::vector<int>& plc = pl.plocal[(size_t)qi];
        plc.assign((size_t)g.nump, -1);
        for (size_t i = 0; i < pg.size(); ++i) plc[(size_t)pg[i]] = (int)i;
    }
    // local meshes; a master point's corner list spans every piece
    for (int qi = 0; qi < npc; ++qi) {
        const int c = pl.pid[(size_t)qi];
        const std::vector<int>& zg = pl.zglob[(size_t)qi];
        const std::vector<int>& sg = pl.sglob[(size_t)qi];
        const std::vector<int>& pg = pl.pglob[(size_t)qi];
        const std::vector<int>& plc = pl.plocal[(size_t)qi];
        HostMesh& h = pl.meshes[(size_t)qi];
        h.numz = (int)zg.size(); h.nums = (int)sg.size(); h.nump = (int)pg.size();
        h.px.resize((size_t)h.nump); h.py.resize((size_t)h.nump);
        for (int i = 0; i < h.nump; ++i) {
            h.px[(size_t)i] = g.px[(size_t)pg[(size_t)i]];
            h.py[(size_t)i] = g.py[(size_t)pg[(size_t)i]];
        }
        h.znump.resize((size_t)h.numz); h.zsbeg.resize((size_t)h.numz + 1);
        std::vector<int> zloc((size_t)g.numz, -1);
        for (int i = 0; i < h.numz; ++i) zloc[(size_t)zg[(size_t)i]] = i;
        int sacc = 0;
        h.maxznump = 0;
        for (int i = 0; i < h.numz; ++i) {
            h.zsbeg[(size_t)i] = sacc;
            h.znump[(size_t)i] = g.znump[(size_t)zg[(size_t)i]];
            h.maxznump = std::max(h.maxznump, h.znump[(size_t)i]);
            sacc += h.znump[(size_t)i];
        }
        h.zsbeg[(size_t)h.numz] = sacc;
        h.mapsp1.resize((size_t)h.nums); h.mapsp2.resize((size_t)h.nums); h.mapsz.resize((size_t)h.nums);
        h.mapss3.resize((size_t)h.nums); h.mapss4.resize((size_t)h.nums);
        for (int i = 0; i < h.nums; ++i) {
            const int s = sg[(size_t)i];
            h.mapsp1[(size_t)i] = plc[(size_t)g.mapsp1[(size_t)s]];
            h.mapsp2[(size_t)i] = plc[(size_t)g.mapsp2[(size_t)s]];
            h.mapsz[(size_t)i] = zloc[(size_t)g.mapsz[(size_t)s]];
            h.mapss3[(size_t)i] = pl.slocal[(size_t)g.mapss3[(size_t)s]];
            h.mapss4[(size_t)i] = pl.slocal[(size_t)g.mapss4[(size_t)s]];
        }
        h.cyclic = true;
        const int nown = pl.nprv[(size_t)qi] + pl.nmstr[(size_t)qi];
        h.pcbeg.assign((size_t)h.nump + 1, 0);
        h.pcside.clear();
        for (int lp = 0; lp < h.nump; ++lp) {
            const int gp = pg[(size_t)lp];
            h.pcbeg[(size_t)lp] = (int)h.pcside.size();
            int cpos = 0;
            for (int k = g.pcbeg[(size_t)gp]; k < g.pcbeg[(size_t)gp + 1]; ++k) {
                const int s = g.pcside[(size_t)k];          // ascending one-piece side order
                const bool mine = (zpc[(size_t)g.mapsz[(size_t)s]] == c);
                if (mine) {
                    h.pcside.push_back(pl.slocal[(size_t)s]);
                } else if (lp >= pl.nprv[(size_t)qi] && lp < nown) {
                    pl.recvs.push_back(Recv{qi, lp, cpos, (int)h.pcside.size(), s});
                    h.pcside.push_back(-1);
                    pl.nrecv[(size_t)qi] += 1;
                } else {
                    continue;
                }
                ++cpos;
            }
        }
        h.pcbeg[(size_t)h.nump] = (int)h.pcside.size();
        if (h.maxznump > PENNANT_MAX_ZSIDES) { err = "zones with more than 8 sides are not supported"; return false; }
    }
    return true;
}

}  // namespace

extern "C" {

pennant_pieces* pennant_pieces_create(const char* deck_text, int numpcs, char* errbuf, int errlen) {
    pennant_pieces* pp = new pennant_pieces();
    auto fail = [&](const std::string& msg) -> pennant_pieces* {
        if (errbuf && errlen > 0) snprintf(errbuf, errlen, "%s", msg.c_str());
        pennant_pieces_destroy(pp);
        return nullptr;
    };
    std::string err;
    if (numpcs < 1) return fail("numpcs must be at least 1");
    if (!read_params(deck_text, pp->prm, err)) return fail(err);
    const Params& q = pp->prm;
    HostMesh g;
    build_mesh(q, g);
    pp->numz = g.numz; pp->nums = g.nums; pp->nump = g.nump;
    Plan pl;
    if (!plan_pieces(q, g, numpcs, pl, err)) return fail(err);
    pp->numpcs = numpcs; pp->numpcx = pl.numpcx; pp->numpcy = pl.numpcy;
    pp->pid = pl.pid;
    pp->zglob = pl.zglob; pp->sglob = pl.sglob; pp->pglob = pl.pglob;
    pp->nprv = pl.nprv; pp->nmstr = pl.nmstr; pp->nghost = pl.nghost; pp->nrecv = pl.nrecv;
    const int npc = (int)pl.pid.size();
    const std::vector<int>& own = pl.own;
    const std::vector<int>& idx_of_piece = pl.idx_of_piece;
    const std::vector<int>& slocal = pl.slocal;
    const std::vector<int>& spiece = pl.spiece;
    const std::vector<std::vector<int>>& plocal = pl.plocal;
    std::vector<HostMesh>& meshes = pl.meshes;
    const std::vector<Recv>& recvs = pl.recvs;
    // the pieces, each initialised as the one-piece run initialises its zones and points
    for (int qi = 0; qi < npc; ++qi) {
        const int nown = pp->nprv[(size_t)qi] + pp->nmstr[(size_t)qi];
        pennant_problem* pb = create_problem(q, std::move(meshes[(size_t)qi]), nown, err);
        if (!pb) return fail("piece " + std::to_string(pp->pid[(size_t)qi]) + ": " + err);
        pb->tune.pipeline = 0;
        pp->pc.push_back(pb);
    }
    // fork and join events of the piece cycle's phases
    pp->ev_fork.assign(3, nullptr);
    pp->ev_join.assign((size_t)3 * npc, nullptr);
    for (hipEvent_t& e : pp->ev_fork)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return fail("hipEventCreate failed");
    for (hipEvent_t& e : pp->ev_join)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return fail("hipEventCreate failed");
    // corner exchange: the slot a ghost's corner is written to, and the
    // owner's slot for it (both as corner_slot places them)
    std::vector<XEnt> cx;
    cx.reserve(recvs.size());
    for (const Recv& r : recvs) {
        const int sp = spiece[(size_t)r.gside], ls = slocal[(size_t)r.gside];
        const HostMesh& hs = pp->pc[(size_t)sp]->hm;
        const int lp = hs.mapsp1[(size_t)ls];       // the ghost point in piece sp
        int src = -1;
        for (int k = hs.pcbeg[(size_t)lp]; k < hs.pcbeg[(size_t)lp + 1]; ++k)
            if (hs.pcside[(size_t)k] == ls) src = (int)corner_slot(hs.nump, lp, k - hs.pcbeg[(size_t)lp], k);
        if (src < 0) return fail("corner exchange: a ghost corner is missing from its piece");
        const HostMesh& hd = pp->pc[(size_t)r.owner]->hm;
        cx.push_back(XEnt{sp, src, r.owner, (int)corner_slot(hd.nump, r.lp, r.c, r.k)});
    }
    // halo: every ghost point from its owner
    std::vector<XEnt> halo;
    for (int qi = 0; qi < npc; ++qi) {
        const int nown = pp->nprv[(size_t)qi] + pp->nmstr[(size_t)qi];
        const std::vector<int>& pg = pp->pglob[(size_t)qi];
        for (int lp = nown; lp < (int)pg.size(); ++lp) {
            const int gp = pg[(size_t)lp];
            const int o = idx_of_piece[(size_t)own[(size_t)gp]];
            if (o < 0) return fail("halo: a ghost point's owner piece is empty");
            const int src = plocal[(size_t)o][(size_t)gp];
            if (src < 0 || src >= pp->nprv[(size_t)o] + pp->nmstr[(size_t)o])
                return fail("halo: a ghost point is not among its owner's points");
            halo.push_back(XEnt{o, src, qi, lp});
        }
    }
    std::vector<PieceTab> tab((size_t)npc);
    for (int qi = 0; qi < npc; ++qi) {
        pennant_problem* pb = pp->pc[(size_t)qi];
        PieceTab& t = tab[(size_t)qi];
        t.pxx = pb->f("px", 0); t.pxy = pb->f("px", 1); t.pux = pb->f("pu", 0); t.puy = pb->f("pu", 1);
        t.vs2 = pb->vs2; t.pxpx = pb->pxpx; t.cfx = pb->cfx; t.cfy = pb->cfy; t.cmass = pb->cmass;
        t.fcrn = pb->fcrn;
    }
    pp->tab = upload_to(pp, tab);
    pp->cx = upload_to(pp, cx);
    pp->ncx = (int)cx.size();
    pp->halo = upload_to(pp, halo);
    pp->nhalo = (int)halo.size();
    int nruns = 0;
    for (pennant_problem* pb : pp->pc) nruns += (int)pb->zruns.size() / 3;
    pp->npart = nruns * PN_MAX_GRID;
    pp->part_d = upload_to(pp, std::vector<double>((size_t)2 * pp->npart, 0.));
    pp->part_i = upload_to(pp, std::vector<int>((size_t)2 * pp->npart, 0));
    if (!pp->tab || !pp->cx || !pp->halo || !pp->part_d || !pp->part_i) return fail("device allocation failed");
    if (hipDeviceSynchronize() != hipSuccess) return fail("piece set-up failed");
    return pp;
}

// The partition alone, on the host (no device): out[0] numpcx, out[1]
// numpcy, out[2] the number of non-empty pieces n, then per piece 7 values:
// piece number, zones, sides, private, master, ghost and received corners.
// Returns the number of int64 values written (3 + 7 n), or PENNANT_EINVAL
// with a message in errbuf when the deck is invalid or out is too small.
int pennant_pieces_plan(const char* deck_text, int numpcs, int64_t* out, int64_t outlen, char* errbuf,
                        int errlen) {
    Params q;
    std::string err;
    auto fail = [&](const std::string& msg) {
        if (errbuf && errlen > 0) snprintf(errbuf, errlen, "%s", msg.c_str());
        return PENNANT_EINVAL;
    };
    if (numpcs < 1) return fail("numpcs must be at least 1");
    if (!read_params(deck_text, q, err)) return fail(err);
    HostMesh g;
    build_mesh(q, g);
    Plan pl;
    if (!plan_pieces(q, g, numpcs, pl, err)) return fail(err);
    const int64_t n = (int64_t)pl.pid.size();
    if (outlen < 3 + 7 * n) return fail("output array too small");
    out[0] = pl.numpcx; out[1] = pl.numpcy; out[2] = n;
    for (int64_t i = 0; i < n; ++i) {
        int64_t* o = out + 3 + 7 * i;
        o[0] = pl.pid[(size_t)i];
        o[1] = (int64_t)pl.zglob[(size_t)i].size();
        o[2] = (int64_t)pl.sglob[(size_t)i].size();
        o[3] = pl.nprv[(size_t)i];
        o[4] = pl.nmstr[(size_t)i];
        o[5] = pl.nghost[(size_t)i];
        o[6] = pl.nrecv[(size_t)i];
    }
    return (int)(3 + 7 * n);
}

void pennant_pieces_destroy(pennant_pieces* pp) {
    if (!pp) return;
    if (!pp->pc.empty()) (void)hipStreamSynchronize(pp->pc[0]->stream);
    if (pp->gexec) (void)hipGraphExecDestroy(pp->gexec);
    for (hipEvent_t e : pp->ev_fork)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : pp->ev_join)
        if (e) (void)hipEventDestroy(e);
    for (pennant_problem* pb : pp->pc) pennant_destroy(pb);
    for (void* p : pp->allocs) (void)hipFree(p);
    delete pp;
}

const char* pennant_pieces_error(pennant_pieces* pp) { return pp->err.c_str(); }

int pennant_pieces_layout(pennant_pieces* pp, int piece, int64_t* out8) {
    if (piece < 0 || piece >= (int)pp->pc.size()) return PENNANT_EINVAL;
    const pennant_problem* pb = pp->pc[(size_t)piece];
    out8[0] = pp->pid[(size_t)piece];
    out8[1] = pb->hm.numz;
    out8[2] = pb->hm.nums;
    out8[3] = pp->nprv[(size_t)piece];
    out8[4] = pp->nmstr[(size_t)piece];
    out8[5] = pp->nghost[(size_t)piece];
    out8[6] = pp->nrecv[(size_t)piece];
    out8[7] = (int64_t)pp->numpcx * 65536 + pp->numpcy;
    return 0;
}

int pennant_pieces_count(pennant_pieces* pp) { return (int)pp->pc.size(); }

int pennant_pieces_run(pennant_pieces* pp, int ncycles, int mode) {
    const int form = fused_form_of(mode);
    if (form < 0) return PENNANT_EINVAL;
    int rc = sync_ctl(pp);
    if (rc) return -rc;
    const Ctl& c = *pp->pc[0]->ctl_host;
    const int c0 = c.cycle;
    int left = ncycles;
    while (left > 0 && !c.done) {
        const int k = left < 64 ? left : 64;
        if ((rc = launch_pieces_cycles(pp, k, form))) return rc > 0 ? -rc : rc;
        if ((rc = sync_ctl(pp))) return -rc;
        left -= k;
    }
    if (c.err_cycle >= 0) {
        char b[160];
        snprintf(b, sizeof b, "Error: %d negative side volumes on cycle %d",
                 c.err_pred ? c.numsbad_pred : c.numsbad_corr, c.err_cycle);
        pp->err = b;
        return PENNANT_EBADSIDES;
    }
    return c.cycle - c0;
}

// ncycles cycles between two events on piece 0's stream (no host wait in
// between; runs of 16 cycles replay the captured graph); *total_ms their time.
int pennant_pieces_time_cycles(pennant_pieces* pp, int ncycles, int mode, double* total_ms) {
    const int form = fused_form_of(mode);
    if (form < 0 || pp->pc.empty()) return PENNANT_EINVAL;
    hipStream_t s = pp->pc[0]->stream;
    int rc = 0;
    if (form == FORM_FP64 && (rc = pieces_materialize_zre(pp))) return rc;
    // the graph is built outside the timed interval
    if (ncycles >= PN_PIECES_GRAPH_CYCLES && (pp->graph_form != form || !pp->gexec) &&
        (rc = build_pieces_graph(pp, form)))
        return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if ((rc = (int)hipEventCreate(&e0)) || (rc = (int)hipEventCreate(&e1))) return rc;
    rc = (int)hipEventRecord(e0, s);
    if (!rc) rc = launch_pieces_cycles(pp, ncycles, form);
    if (!rc) rc = (int)hipEventRecord(e1, s);
    if (!rc) rc = (int)hipEventSynchronize(e1);
    float ms = 0.f;
    if (!rc) rc = (int)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (total_ms) *total_ms = ms;
    return rc;
}

int pennant_pieces_scalars(pennant_pieces* pp, double* o) {
    const int rc = sync_ctl(pp);
    if (rc) return rc;
    const Ctl& c = *pp->pc[0]->ctl_host;
    o[0] = c.time; o[1] = c.dtlast; o[2] = c.dtrec; o[3] = c.cycle; o[4] = c.dtnew;
    o[5] = c.dvovmax; o[6] = c.numsbad_pred; o[7] = c.numsbad_corr; o[8] = c.dt; o[9] = c.done;
    return 0;
}

int pennant_pieces_field_info(pennant_pieces* pp, const char* name, int64_t* count, int* ncomp) {
    if (pp->pc.empty()) return PENNANT_EINVAL;
    auto it = pp->pc[0]->fields.find(name);
    if (it == pp->pc[0]->fields.end()) return PENNANT_EINVAL;
    const int kind = it->second.kind;
    *count = kind == 0 ? pp->nump : (kind == 1 ? pp->numz : pp->nums);
    *ncomp = it->second.ncomp;
    return 0;
}

// A zone, side or point field in the one-piece numbering: zones and sides
// from their piece, points from their owner.
int pennant_pieces_get_field(pennant_pieces* pp, const char* name, double* out) {
    int rc = pieces_materialize_zre(pp);
    if (rc) return rc;
    if (pp->pc.empty()) return PENNANT_EINVAL;
    auto it = pp->pc[0]->fields.find(name);
    if (it == pp->pc[0]->fields.end()) return PENNANT_EINVAL;
    const int kind = it->second.kind, ncomp = it->second.ncomp;
    if ((rc = (int)hipStreamSynchronize(pp->pc[0]->stream))) return rc;
    for (size_t qi = 0; qi < pp->pc.size(); ++qi) {
        pennant_problem* pb = pp->pc[qi];
        const std::vector<int>& map = kind == 0 ? pp->pglob[qi] : (kind == 1 ? pp->zglob[qi] : pp->sglob[qi]);
        const size_t n = kind == 0 ? (size_t)(pp->nprv[qi] + pp->nmstr[qi]) : map.size();
        std::vector<double> tmp(n);
        for (int c = 0; c < ncomp; ++c) {
            if (n && (rc = (int)hipMemcpy(tmp.data(), pb->fields.at(name).c[c], n * sizeof(double),
                                          hipMemcpyDeviceToHost)))
                return rc;
            for (size_t i = 0; i < n; ++i) out[(size_t)map[i] * ncomp + c] = tmp[i];
        }
    }
    return 0;
}

}  // extern "C"
