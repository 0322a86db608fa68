// Batched problems: many independent PENNANT problems of uneven sizes
// advanced together, one launch per phase (fused.hip, k_*_batch).
//
// Every problem is created and initialised by the single-problem driver
// (pennant_create: deck parsing, mesh generation, geometry, state, first dt),
// then its state, maps and control block are copied into one arena with
// arena-wide indices, and the problem is freed.  Each problem keeps its own
// control block (dt, time, cycle, stop flag) and parameters, so the batch
// computes for every problem exactly what its own run computes.
//
// Cycles are launched through a captured HIP graph of PN_GRAPH_CYCLES cycles
// (four kernels each, no memset or host work inside), replayed as often as
// needed.  The autotuner times grid shapes on a snapshot of the state and
// restores it afterwards, so tuning never advances the problems.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "common.h"
#include "problem.h"

using namespace pn;

#define PN_GRAPH_CYCLES 16

namespace {

#define BCK(x) do { const int e_ = (int)(x); if (e_ != 0) return e_; } while (0)

int batch_form(int mode) {
    switch (mode) {
        case PENNANT_MODE_FUSED: return FORM_FP64;
        case PENNANT_MODE_MIXED: return FORM_MIXED;
        case PENNANT_MODE_LEARNED: return FORM_LEARNED;
        default: return -1;
    }
}

void drop_graph(pennant_batch* b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
    b->graph_form = -1;
    b->graph_cycles = 0;
}

int build_graph(pennant_batch* b, int form) {
    drop_graph(b);
    hipGraph_t g = nullptr;
    BCK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int k = 0; k < PN_GRAPH_CYCLES && !rc; ++k) rc = batch_cycle(b, form);
    const int rc2 = (int)hipStreamEndCapture(b->stream, &g);
    if (rc || rc2) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : rc2;
    }
    rc = (int)hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (rc) { b->gexec = nullptr; return rc; }
    b->graph_form = form;
    b->graph_cycles = PN_GRAPH_CYCLES;
    return 0;
}

// state arrays a cycle changes (for the autotuner's snapshot)
struct Snap {
    std::vector<std::pair<void*, size_t>> arrays;
    std::vector<void*> copies;
};

std::vector<std::pair<void*, size_t>> state_arrays(pennant_batch* b) {
    const size_t Z = (size_t)b->numz * 8, P = (size_t)b->nump * 8;
    return {{b->pxx, P}, {b->pxy, P}, {b->pux, P}, {b->puy, P}, {b->zr, Z}, {b->ze, Z},
            {b->zetot, Z}, {b->zwrate, Z}, {b->zvol, Z}, {b->zp, Z},
            {b->ctls, sizeof(Ctl) * (size_t)b->nprob},
            // the closure and capture counters, so the timing cycles do not count
            {b->cl_counters, 3 * sizeof(unsigned long long)}};
}

int snap_take(pennant_batch* b, Snap& sn) {
    sn.arrays = state_arrays(b);
    for (auto& a : sn.arrays) {
        void* c = nullptr;
        BCK(hipMalloc(&c, a.second));
        sn.copies.push_back(c);
        BCK(hipMemcpyAsync(c, a.first, a.second, hipMemcpyDeviceToDevice, b->stream));
    }
    return (int)hipStreamSynchronize(b->stream);
}

int snap_restore(pennant_batch* b, const Snap& sn) {
    for (size_t i = 0; i < sn.arrays.size(); ++i)
        BCK(hipMemcpyAsync(sn.arrays[i].first, sn.copies[i], sn.arrays[i].second, hipMemcpyDeviceToDevice,
                           b->stream));
    return (int)hipStreamSynchronize(b->stream);
}

void snap_free(Snap& sn) {
    for (void* c : sn.copies) (void)hipFree(c);
    sn.copies.clear();
}

int copy_ctls(pennant_batch* b) {
    BCK(hipMemcpyAsync(b->ctls_host, b->ctls, sizeof(Ctl) * b->nprob, hipMemcpyDeviceToHost, b->stream));
    return (int)hipStreamSynchronize(b->stream);
}

}  // namespace

extern "C" {

pennant_batch* pennant_batch_create(const char* const* decks, int n, char* errbuf, int errlen) {
    auto fail = [&](pennant_batch* b, const std::string& m) -> pennant_batch* {
        if (errbuf && errlen > 0) snprintf(errbuf, (size_t)errlen, "%s", m.c_str());
        pennant_batch_destroy(b);
        return nullptr;
    };
    if (n <= 0 || !decks) return fail(nullptr, "a batch needs at least one deck");
    pennant_batch* b = new pennant_batch();
    b->nprob = n;
    // pass 1: sizes from the host-side mesh generator
    char e[512];
    b->zoff.assign(n + 1, 0);
    b->soff.assign(n + 1, 0);
    b->poff.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        pennant_host_mesh* hm = pennant_host_mesh_create(decks[i], e, sizeof e);
        if (!hm) return fail(b, "deck " + std::to_string(i) + ": " + e);
        int64_t nz = 0, ns = 0, np = 0;
        pennant_host_mesh_get(hm, "znump", nullptr, &nz);
        pennant_host_mesh_get(hm, "mapsp1", nullptr, &ns);
        pennant_host_mesh_get(hm, "px", nullptr, &np);
        pennant_host_mesh_destroy(hm);
        b->zoff[i + 1] = b->zoff[i] + nz;
        b->soff[i + 1] = b->soff[i] + ns;
        b->poff[i + 1] = b->poff[i] + np;
    }
    b->numz = b->zoff[n];
    b->nums = b->soff[n];
    b->nump = b->poff[n];
    if (b->nums >= (int64_t)1 << 31 || b->nump >= (int64_t)1 << 31)
        return fail(b, "batch too large for int32 indices");
    if (hipStreamCreate(&b->stream) != hipSuccess) return fail(b, "hipStreamCreate failed");
    bool ok = true;
    auto al = [&](size_t count, size_t sz) -> void* {
        void* p = nullptr;
        const size_t nb = count * sz > 0 ? count * sz : sz;
        if (hipMalloc(&p, nb) != hipSuccess || hipMemset(p, 0, nb) != hipSuccess) { ok = false; return nullptr; }
        b->allocs.push_back(p);
        return p;
    };
    const size_t Z = b->numz, S = b->nums, P = b->nump;
    if (!corner_slots_fit((int64_t)P, (int64_t)S)) return fail(b, "batch too large for 32-bit corner slots");
    const size_t NC = (size_t)corner_slots((int64_t)P, (int64_t)S);
    b->pxx = (double*)al(P, 8); b->pxy = (double*)al(P, 8); b->pux = (double*)al(P, 8); b->puy = (double*)al(P, 8);
    b->zr = (double*)al(Z, 8); b->ze = (double*)al(Z, 8); b->zetot = (double*)al(Z, 8);
    b->zwrate = (double*)al(Z, 8); b->zvol = (double*)al(Z, 8); b->zp = (double*)al(Z, 8);
    b->zm = (double*)al(Z, 8); b->smf = (double*)al(S, 8); b->fsmf = (float*)al(S, 4);
    b->mapsp1 = (int*)al(S, 4); b->pcbeg = (int*)al(P + 1, 4); b->pcpos = (int*)al(S, 4);
    b->pbc = (unsigned char*)al(P, 1);
    b->cfx = (double*)al(NC, 8); b->cfy = (double*)al(NC, 8); b->cmass = (double*)al(NC, 8);
    b->sfwx = (double*)al(S, 8); b->sfwy = (double*)al(S, 8);
    b->fcrn = (float*)al(3 * NC, 4); b->fsfw = (float*)al(2 * S, 4);
    b->vs2 = (double*)al(2 * P, 8); b->pxpx = (double*)al(P, 8);
    b->ctls = (Ctl*)al((size_t)n, sizeof(Ctl));
    b->probs = (BProb*)al((size_t)n, sizeof(BProb));
    b->cl_counters = (unsigned long long*)al(3, 8);
    if (!ok) return fail(b, "device allocation failed");
    if (hipHostMalloc((void**)&b->ctls_host, sizeof(Ctl) * n) != hipSuccess) return fail(b, "pinned allocation failed");

    // pass 2: each problem through the single-problem driver, then into the arena
    std::vector<int> mapsp1(S), pcbeg(P + 1), pcpos(S);
    std::vector<BUnit> units;
    std::vector<PUnit> punits;
    std::vector<BHeavy> heavy;
    std::vector<BProb> probs(n);
    for (int i = 0; i < n; ++i) {
        pennant_problem* pb = pennant_create(decks[i], e, sizeof e);
        if (!pb) return fail(b, "deck " + std::to_string(i) + ": " + e);
        const HostMesh& h = pb->hm;
        const int64_t zo = b->zoff[i], so = b->soff[i], po = b->poff[i];
        const size_t nz = h.numz, ns = h.nums, np = h.nump;
        struct { double* dst; const double* src; size_t n; } cp[] = {
            {b->pxx + po, pb->f("px", 0), np}, {b->pxy + po, pb->f("px", 1), np},
            {b->pux + po, pb->f("pu", 0), np}, {b->puy + po, pb->f("pu", 1), np},
            {b->zr + zo, pb->f("zr"), nz}, {b->ze + zo, pb->f("ze"), nz},
            {b->zetot + zo, pb->f("zetot"), nz}, {b->zwrate + zo, pb->f("zwrate"), nz},
            {b->zvol + zo, pb->f("zvol"), nz}, {b->zp + zo, pb->f("zp"), nz},
            {b->zm + zo, pb->f("zm"), nz}, {b->smf + so, pb->f("smf"), ns}};
        int rc = (int)hipStreamSynchronize(pb->stream);
        for (auto& c : cp)
            if (!rc) rc = (int)hipMemcpy(c.dst, c.src, c.n * 8, hipMemcpyDeviceToDevice);
        if (!rc) rc = (int)hipMemcpy(b->pbc + po, pb->pbc, np, hipMemcpyDeviceToDevice);
        if (!rc) rc = (int)hipMemcpy(b->ctls + i, pb->ctl, sizeof(Ctl), hipMemcpyDeviceToDevice);
        if (rc) { pennant_destroy(pb); return fail(b, "arena copy failed"); }
        for (size_t s = 0; s < ns; ++s) mapsp1[so + s] = (int)(po + h.mapsp1[s]);
        for (size_t p = 0; p < np; ++p) pcbeg[po + p] = (int)(so + h.pcbeg[p]);
        // corner slots of the arena (its points numbered po + p)
        for (size_t p = 0; p < np; ++p)
            for (int k = h.pcbeg[p]; k < h.pcbeg[p + 1]; ++k)
                pcpos[so + h.pcside[k]] =
                    (int)corner_slot((int64_t)P, po + (int64_t)p, k - h.pcbeg[p], so + k);
        for (size_t p = 0; p < np; ++p)
            if (h.pcbeg[p + 1] - h.pcbeg[p] > PN_HEAVY) heavy.push_back({(int)(po + p), i});
        BProb& q = probs[i];
        const Params& pr = pb->prm;
        q.gamma = pr.gamma; q.ssmin = pr.ssmin; q.alfa = pr.alfa; q.qgamma = pr.qgamma;
        q.q1 = pr.q1; q.q2 = pr.q2; q.cfl = pr.cfl; q.cflv = pr.cflv; q.dtmax = pr.dtmax;
        q.dtfac = pr.dtfac; q.tstop = pr.tstop; q.cstop = pr.cstop; q.nbcs = pb->nbcs;
        for (int k = 0; k < 8; ++k) { q.bcvx[k] = pb->bcvx[k]; q.bcvy[k] = pb->bcvy[k]; }
        // the single-problem set-up's rect detection (every side checked)
        q.rect_nzx = pb->rect_nzx;
        q.rect_m = pb->rect_m;
        q.rect_sh = pb->rect_sh;
        q.so = (int)so;
        q.po = (int)po;
        q.ubeg = (int)units.size();
        for (size_t r = 0; r < pb->zruns.size(); r += 3) {
            const int zb = pb->zruns[r], ze = pb->zruns[r + 1], nsd = pb->zruns[r + 2];
            if (nsd != 3 && nsd != 4) { pennant_destroy(pb); return fail(b, "batched zones must have 3 or 4 sides"); }
            for (int u0 = zb; u0 < ze; u0 += PN_BLOCK) {
                BUnit U;
                std::memset(&U, 0, sizeof U);
                U.zb = (int)(zo + u0);
                U.ze = (int)(zo + (u0 + PN_BLOCK < ze ? u0 + PN_BLOCK : ze));
                U.sb = (int)(so + h.zsbeg[zb] + (int64_t)(u0 - zb) * nsd);
                U.ns = nsd;
                U.pid = i;
                units.push_back(U);
            }
        }
        q.uend = (int)units.size();
        for (size_t p0 = 0; p0 < np; p0 += PN_BLOCK)
            punits.push_back({(int)(po + p0), (int)(po + (p0 + PN_BLOCK < np ? p0 + PN_BLOCK : np)), i, 0});
        pennant_destroy(pb);
    }
    pcbeg[P] = (int)S;
    b->nunits = (int)units.size();
    b->npunits = (int)punits.size();
    b->nheavy = (int)heavy.size();
    b->units = (BUnit*)al(units.size(), sizeof(BUnit));
    b->punits = (PUnit*)al(punits.size(), sizeof(PUnit));
    b->heavy = (BHeavy*)al(heavy.size(), sizeof(BHeavy));
    b->part_d = (double*)al(2 * units.size(), 8);
    b->part_i = (int*)al(2 * units.size(), 4);
    if (!ok) return fail(b, "device allocation failed");
    int rc = (int)hipMemcpy(b->mapsp1, mapsp1.data(), S * 4, hipMemcpyHostToDevice);
    if (!rc) rc = (int)hipMemcpy(b->pcbeg, pcbeg.data(), (P + 1) * 4, hipMemcpyHostToDevice);
    if (!rc) rc = (int)hipMemcpy(b->pcpos, pcpos.data(), S * 4, hipMemcpyHostToDevice);
    if (!rc) rc = (int)hipMemcpy(b->units, units.data(), units.size() * sizeof(BUnit), hipMemcpyHostToDevice);
    if (!rc) rc = (int)hipMemcpy(b->punits, punits.data(), punits.size() * sizeof(PUnit), hipMemcpyHostToDevice);
    if (!rc && !heavy.empty())
        rc = (int)hipMemcpy(b->heavy, heavy.data(), heavy.size() * sizeof(BHeavy), hipMemcpyHostToDevice);
    if (!rc) rc = (int)hipMemcpy(b->probs, probs.data(), n * sizeof(BProb), hipMemcpyHostToDevice);
    if (!rc) rc = launch_to_float(b->smf, b->fsmf, (int64_t)S, b->stream);
    if (!rc) rc = (int)hipStreamSynchronize(b->stream);
    if (rc) return fail(b, "batch upload failed");
    return b;
}

void pennant_batch_destroy(pennant_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (void* p : b->allocs) (void)hipFree(p);
    if (b->ctls_host) (void)hipHostFree(b->ctls_host);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int pennant_batch_count(pennant_batch* b) { return b->nprob; }

int pennant_batch_sizes(pennant_batch* b, int i, int64_t* out3) {
    if (i < 0 || i >= b->nprob) return PENNANT_EINVAL;
    out3[0] = b->poff[i + 1] - b->poff[i];
    out3[1] = b->zoff[i + 1] - b->zoff[i];
    out3[2] = b->soff[i + 1] - b->soff[i];
    return 0;
}

const char* pennant_batch_error(pennant_batch* b) { return b->err.c_str(); }

int pennant_batch_launch_cycles(pennant_batch* b, int ncycles, int mode) {
    const int form = batch_form(mode);
    if (form < 0 || ncycles < 0) return PENNANT_EINVAL;
    if (form == FORM_LEARNED && !b->closure) return PENNANT_EINVAL;
    // the FP64 form reads zr and ze; mixed cycles leave them derived
    if (form == FORM_FP64) BCK(batch_materialize_zre(b));
    else b->zre_stale = true;
    int left = ncycles;
    if (b->tune.graph && left >= PN_GRAPH_CYCLES) {
        if (b->graph_form != form || !b->gexec) BCK(build_graph(b, form));
        for (; left >= b->graph_cycles; left -= b->graph_cycles) BCK(hipGraphLaunch(b->gexec, b->stream));
    }
    for (; left > 0; --left) BCK(batch_cycle(b, form));
    return 0;
}

int pennant_batch_sync(pennant_batch* b) {
    BCK(copy_ctls(b));
    for (int i = 0; i < b->nprob; ++i) {
        const Ctl& c = b->ctls_host[i];
        if (c.err_cycle >= 0) {
            char m[192];
            snprintf(m, sizeof m, "Error: %d negative side volumes on cycle %d (problem %d)",
                     c.err_pred ? c.numsbad_pred : c.numsbad_corr, c.err_cycle, i);
            b->err = m;
            return PENNANT_EBADSIDES;
        }
    }
    return 0;
}

// number of problems still running (0: all stopped)
int pennant_batch_active(pennant_batch* b) {
    BCK(copy_ctls(b));
    int a = 0;
    for (int i = 0; i < b->nprob; ++i) a += b->ctls_host[i].done ? 0 : 1;
    return a;
}

int pennant_batch_time_cycles(pennant_batch* b, int ncycles, int mode, double* total_ms) {
    if (ncycles < 1) return PENNANT_EINVAL;
    const int form = batch_form(mode);
    if (form < 0) return PENNANT_EINVAL;
    // build the graph outside the timed region
    if (b->tune.graph && ncycles >= PN_GRAPH_CYCLES && (b->graph_form != form || !b->gexec))
        BCK(build_graph(b, form));
    hipEvent_t e0, e1;
    BCK(hipEventCreate(&e0));
    BCK(hipEventCreate(&e1));
    int rc = (int)hipEventRecord(e0, b->stream);
    if (!rc) rc = pennant_batch_launch_cycles(b, ncycles, mode);
    if (!rc) rc = (int)hipEventRecord(e1, b->stream);
    if (!rc) rc = (int)hipEventSynchronize(e1);
    float ms = 0.f;
    if (!rc) rc = (int)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    *total_ms = ms;
    return pennant_batch_sync(b);
}

int pennant_batch_scalars(pennant_batch* b, int i, double* o) {
    if (i < 0 || i >= b->nprob) return PENNANT_EINVAL;
    BCK(copy_ctls(b));
    const Ctl& c = b->ctls_host[i];
    o[0] = c.time; o[1] = c.dtlast; o[2] = c.dtrec; o[3] = c.cycle; o[4] = c.dtnew;
    o[5] = c.dvovmax; o[6] = c.numsbad_pred; o[7] = c.numsbad_corr; o[8] = c.dt; o[9] = c.done;
    return 0;
}

int pennant_batch_get_field(pennant_batch* b, int i, const char* name, double* out) {
    if (i < 0 || i >= b->nprob) return PENNANT_EINVAL;
    BCK(batch_materialize_zre(b));
    const int64_t zo = b->zoff[i], nz = b->zoff[i + 1] - zo;
    const int64_t po = b->poff[i], np = b->poff[i + 1] - po;
    const std::string f(name);
    const double* c[2] = {nullptr, nullptr};
    int64_t off = zo, cnt = nz;
    int ncomp = 1;
    if (f == "px") { c[0] = b->pxx; c[1] = b->pxy; ncomp = 2; off = po; cnt = np; }
    else if (f == "pu") { c[0] = b->pux; c[1] = b->puy; ncomp = 2; off = po; cnt = np; }
    else if (f == "zr") c[0] = b->zr;
    else if (f == "ze") c[0] = b->ze;
    else if (f == "zp") c[0] = b->zp;
    else if (f == "zetot") c[0] = b->zetot;
    else if (f == "zwrate") c[0] = b->zwrate;
    else if (f == "zvol") c[0] = b->zvol;
    else if (f == "zm") c[0] = b->zm;
    else return PENNANT_EINVAL;
    BCK(hipStreamSynchronize(b->stream));
    std::vector<double> tmp((size_t)cnt);
    for (int k = 0; k < ncomp; ++k) {
        BCK(hipMemcpy(tmp.data(), c[k] + off, (size_t)cnt * 8, hipMemcpyDeviceToHost));
        for (int64_t j = 0; j < cnt; ++j) out[j * ncomp + k] = tmp[(size_t)j];
    }
    return 0;
}

const void* pennant_mlp_device_pack(pennant_mlp* m);

int pennant_batch_set_closure(pennant_batch* b, pennant_mlp* m) {
    b->closure = m ? pennant_mlp_device_pack(m) : nullptr;
    drop_graph(b);
    return 0;
}

int pennant_batch_closure_stats(pennant_batch* b, int64_t* out2) {
    unsigned long long c[3];
    BCK(hipMemcpyAsync(c, b->cl_counters, sizeof c, hipMemcpyDeviceToHost, b->stream));
    BCK(hipStreamSynchronize(b->stream));
    out2[0] = (int64_t)c[1];
    out2[1] = (int64_t)c[2];
    return 0;
}

int pennant_batch_set_capture(pennant_batch* b, float* dev_buf, int64_t cap) {
    if (cap < 0 || (dev_buf == nullptr) != (cap == 0)) return PENNANT_EINVAL;
    BCK(hipStreamSynchronize(b->stream));
    b->cap_buf = dev_buf;
    b->cap = cap;
    drop_graph(b);
    const unsigned long long zero = 0;
    return (int)hipMemcpy(b->cl_counters, &zero, sizeof zero, hipMemcpyHostToDevice);
}

int pennant_batch_capture_count(pennant_batch* b, int64_t* count) {
    unsigned long long c = 0;
    BCK(hipMemcpyAsync(&c, b->cl_counters, sizeof c, hipMemcpyDeviceToHost, b->stream));
    BCK(hipStreamSynchronize(b->stream));
    *count = (int64_t)c;
    return 0;
}

int pennant_batch_set_tuning(pennant_batch* b, const char* key, int value) {
    if (!strcmp(key, "max_blocks")) {
        if (value < 1 || value > 65535) return PENNANT_EINVAL;
        b->tune.max_blocks = value;
    } else if (!strcmp(key, "xcd")) {
        if (value != 0 && value != 1) return PENNANT_EINVAL;
        b->tune.xcd = value;
    } else if (!strcmp(key, "graph")) {
        if (value != 0 && value != 1) return PENNANT_EINVAL;
        b->tune.graph = value;
    } else if (!strcmp(key, "rect_maps")) {
        if (value != 0 && value != 1) return PENNANT_EINVAL;
        b->tune.rect_maps = value;
    } else {
        return PENNANT_EINVAL;
    }
    drop_graph(b);
    return 0;
}

int pennant_batch_get_tuning(pennant_batch* b, const char* key) {
    if (!strcmp(key, "max_blocks")) return b->tune.max_blocks;
    if (!strcmp(key, "xcd")) return b->tune.xcd;
    if (!strcmp(key, "graph")) return b->tune.graph;
    if (!strcmp(key, "rect_maps")) return b->tune.rect_maps;
    return PENNANT_EINVAL;
}

int pennant_batch_autotune(pennant_batch* b, int mode, int cycles, double* best_ms) {
    const int form = batch_form(mode);
    if (form < 0 || cycles < 1) return PENNANT_EINVAL;
    if (form == FORM_LEARNED && !b->closure) return PENNANT_EINVAL;
    Snap sn;
    int rc = snap_take(b, sn);
    const Tuning keep = b->tune;
    Tuning best = keep;
    double bms = 1e300;
    const int grids[] = {512, 1024, 2048, 4096, 8192, 16384};
    for (int xcd = 0; xcd <= 1 && !rc; ++xcd) {
        for (int g : grids) {
            if (rc) break;
            b->tune = keep;
            b->tune.max_blocks = g;
            b->tune.xcd = xcd;
            b->tune.graph = 0;
            drop_graph(b);
            rc = snap_restore(b, sn);
            b->zre_stale = true;
            // one untimed cycle, then the timed ones
            if (!rc) rc = pennant_batch_launch_cycles(b, 1, mode);
            double ms = 0.;
            if (!rc) rc = pennant_batch_time_cycles(b, cycles, mode, &ms);
            if (rc == PENNANT_EBADSIDES) rc = 0;
            if (!rc && ms < bms) { bms = ms; best = b->tune; }
        }
    }
    best.graph = keep.graph;
    b->tune = best;
    drop_graph(b);
    const int rc2 = snap_restore(b, sn);
    b->zre_stale = true;
    snap_free(sn);
    if (best_ms) *best_ms = bms;
    return rc ? rc : rc2;
}

// The ensemble's problem parameters: for k = 0..count-1 (even k Sedov on a
// rect mesh, odd k Noh on a pie mesh), n = round(exp(U(ln 32, ln 512)))
// zones per side, then the Sedov blast energy U[0.1, 1.0] or the Noh inflow
// speed U[0.5, 2.0], drawn in that order from std::mt19937_64(seed).
int pennant_ensemble_params(uint64_t seed, int count, int32_t* n, double* param) {
    if (count < 0) return PENNANT_EINVAL;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> lsz(std::log(32.0), std::log(512.0));
    std::uniform_real_distribution<double> en(0.1, 1.0), sp(0.5, 2.0);
    for (int k = 0; k < count; ++k) {
        n[k] = (int32_t)std::lround(std::exp(lsz(rng)));
        param[k] = (k % 2 == 0) ? en(rng) : sp(rng);
    }
    return 0;
}

}  // extern "C"
