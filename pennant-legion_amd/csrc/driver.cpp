// Whole-problem driver: the host side above the task entry points.
//
// Mirrors the reference's host structure -- InputFile (src/InputFile.cc),
// GenMesh one-piece numbering (src/GenMesh.cc), Mesh::init geometry
// (src/Mesh.cc:616-623), HydroBC point lists (src/HydroBC.cc:63-144),
// Hydro::init (src/Hydro.cc:241-339) and Driver::run's cycle loop
// (src/Driver.cc:102-187) -- but keeps every field resident in HBM as SoA
// arrays and keeps dt, time and the stop test on the device.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>

#include "problem.h"

using namespace pn;

namespace {

// ---------------------------------------------------------------------------
// .pnt deck, src/InputFile.cc:24-104.  Like the reference's getline/eof loop,
// a final line with no terminating newline is not read.
struct Deck {
    std::map<std::string, std::string> kv;
    std::string err;
    bool parse(const char* text) {
        const std::string all(text);
        size_t pos = 0;
        while (true) {
            const size_t nl = all.find('\n', pos);
            if (nl == std::string::npos) break;
            std::istringstream ls(all.substr(pos, nl - pos));
            pos = nl + 1;
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
    template <class T> T get(const std::string& k, T d) const {
        auto it = kv.find(k);
        if (it == kv.end()) return d;
        std::istringstream s(it->second);
        T v = T();
        s >> v;
        return v;
    }
    std::vector<double> list(const std::string& k) const {
        std::vector<double> out;
        auto it = kv.find(k);
        if (it == kv.end()) return out;
        std::istringstream s(it->second);
        double v;
        while (s >> v) out.push_back(v);
        return out;
    }
};

}  // namespace

namespace pn {

bool read_params(const char* text, Params& q, std::string& err) {
    Deck d;
    if (!d.parse(text)) { err = d.err; return false; }
    // Driver ctor, src/Driver.cc:78-87
    q.cstop = d.get<int>("cstop", 999999);
    q.tstop = d.get<double>("tstop", 1.e99);
    if (q.cstop == 999999 && q.tstop == 1.e99) { err = "Must specify either cstop or tstop"; return false; }
    q.dtmax = d.get<double>("dtmax", 1.e99);
    q.dtinit = d.get<double>("dtinit", 1.e99);
    q.dtfac = d.get<double>("dtfac", 1.2);
    q.dtreport = d.get<int>("dtreport", 10);
    // Mesh ctor, src/Mesh.cc:216-223
    q.subregion = d.list("subregion");
    if (!q.subregion.empty() && q.subregion.size() != 4) { err = "subregion must have 4 entries"; return false; }
    // GenMesh ctor, src/GenMesh.cc:87-129
    q.meshtype = d.get<std::string>("meshtype", "");
    if (q.meshtype.empty()) { err = "must specify meshtype"; return false; }
    if (q.meshtype != "pie" && q.meshtype != "rect" && q.meshtype != "hex") { err = "invalid meshtype " + q.meshtype; return false; }
    const std::vector<double> mp = d.list("meshparams");
    if (mp.empty()) { err = "must specify meshparams"; return false; }
    if (mp.size() > 4) { err = "meshparams must have <= 4 values"; return false; }
    q.nzx = (long)mp[0];
    q.nzy = mp.size() >= 2 ? (long)mp[1] : q.nzx;
    if (q.meshtype != "pie") q.lenx = mp.size() >= 3 ? mp[2] : 1.0;
    else q.lenx = (mp.size() >= 3 ? mp[2] : 90.0) * M_PI / 180.0;
    q.leny = mp.size() >= 4 ? mp[3] : 1.0;
    if (q.nzx <= 0 || q.nzy <= 0 || q.lenx <= 0. || q.leny <= 0.) { err = "meshparams values must be positive"; return false; }
    if (q.meshtype == "pie" && q.lenx >= 2. * M_PI) { err = "meshparams theta must be < 360"; return false; }
    if (q.meshtype == "hex") { err = "meshtype hex is not supported (its generator is assert(false) in the reference, src/GenMesh.cc:875-896)"; return false; }
    // Hydro ctor, src/Hydro.cc:206-214; PolyGas src/PolyGas.cc:58-59;
    // TTS src/TTS.cc:46-47 (same ssmin key); QCS src/QCS.cc:84-86
    q.cfl = d.get<double>("cfl", 0.6);
    q.cflv = d.get<double>("cflv", 0.1);
    q.rinit = d.get<double>("rinit", 1.);
    q.einit = d.get<double>("einit", 0.);
    q.rinitsub = d.get<double>("rinitsub", 1.);
    q.einitsub = d.get<double>("einitsub", 0.);
    q.uinitradial = d.get<double>("uinitradial", 0.);
    q.bcx = d.list("bcx");
    q.bcy = d.list("bcy");
    q.gamma = d.get<double>("gamma", 5. / 3.);
    q.ssmin = d.get<double>("ssmin", 0.);
    q.alfa = d.get<double>("alfa", 0.5);
    q.qgamma = d.get<double>("qgamma", 5. / 3.);
    q.q1 = d.get<double>("q1", 0.);
    q.q2 = d.get<double>("q2", 2.);
    return true;
}

// ---------------------------------------------------------------------------
// One-piece mesh.  A zone's sides are generated as a cyclic run, so the side
// maps are built zone by zone: side k of zone z joins corner k to corner
// k+1 of the zone's CCW corner list.  The resulting numbering is the
// reference's (src/GenMesh.cc:987-1282 with numpcx = numpcy = 1).

void build_mesh(const Params& q, HostMesh& h) {
    const long nzx = q.nzx, nzy = q.nzy, npx = nzx + 1;
    const bool pie = (q.meshtype == "pie");
    h.numz = (int)(nzx * nzy);
    h.nump = pie ? (int)(npx * nzy + 1) : (int)(npx * (nzy + 1));
    h.px.resize(h.nump);
    h.py.resize(h.nump);
    if (!pie) {
        const double dx = q.lenx / (double)nzx, dy = q.leny / (double)nzy;
        for (long j = 0; j <= nzy; ++j)
            for (long i = 0; i <= nzx; ++i) {
                h.px[j * npx + i] = dx * double(i);
                h.py[j * npx + i] = dy * double(j);
            }
    } else {
        const double dth = q.lenx / (double)nzx, dr = q.leny / (double)nzy;
        h.px[0] = 0.;
        h.py[0] = 0.;
        for (long j = 1; j <= nzy; ++j)
            for (long i = 0; i <= nzx; ++i) {
                const double th = dth * (double)(nzx - i);
                const double r = dr * (double)j;
                // one glibc sincos call, as the reference's g++ -O2 build
                // evaluates cos(th), sin(th) (src/GenMesh.cc:866-867)
                double sn, cs;
                sincos(th, &sn, &cs);
                h.px[(j - 1) * npx + i + 1] = r * cs;
                h.py[(j - 1) * npx + i + 1] = r * sn;
            }
    }
    h.znump.resize(h.numz);
    h.zsbeg.resize(h.numz + 1);
    std::vector<int> corners;   // CCW corner list of every zone, concatenated
    corners.reserve(4 * (size_t)h.numz);
    for (long zy = 0; zy < nzy; ++zy)
        for (long zxi = 0; zxi < nzx; ++zxi) {
            const long z = zy * nzx + zxi;
            h.zsbeg[z] = (int)corners.size();
            if (!pie) {
                const int p0 = (int)(zy * npx + zxi);
                corners.insert(corners.end(), {p0, p0 + 1, p0 + 1 + (int)npx, p0 + (int)npx});
            } else if (zy == 0) {
                const int p0 = (int)(zxi + 1);
                corners.insert(corners.end(), {0, p0 + 1, p0});
            } else {
                const int p0 = (int)((zy - 1) * npx + zxi + 1);
                corners.insert(corners.end(), {p0, p0 + 1, p0 + (int)npx + 1, p0 + (int)npx});
            }
            h.znump[z] = (int)corners.size() - h.zsbeg[z];
        }
    h.zsbeg[h.numz] = (int)corners.size();
    h.nums = (int)corners.size();
    h.mapsp1.resize(h.nums); h.mapsp2.resize(h.nums); h.mapsz.resize(h.nums);
    h.mapss3.resize(h.nums); h.mapss4.resize(h.nums);
    h.maxznump = 0;
    for (int z = 0; z < h.numz; ++z) {
        const int b = h.zsbeg[z], n = h.znump[z];
        if (n > h.maxznump) h.maxznump = n;
        for (int k = 0; k < n; ++k) {
            const int s = b + k;
            h.mapsp1[s] = corners[s];
            h.mapsp2[s] = corners[b + (k + 1) % n];
            h.mapsz[s] = z;
            h.mapss3[s] = b + (k + n - 1) % n;
            h.mapss4[s] = b + (k + 1) % n;
        }
    }
    h.cyclic = true;
    // point-corner CSR: sides grouped by mapsp1, ascending side order
    h.pcbeg.assign(h.nump + 1, 0);
    for (int s = 0; s < h.nums; ++s) h.pcbeg[h.mapsp1[s] + 1]++;
    for (int p = 0; p < h.nump; ++p) h.pcbeg[p + 1] += h.pcbeg[p];
    h.pcside.resize(h.nums);
    std::vector<int> fill(h.pcbeg.begin(), h.pcbeg.end() - 1);
    for (int s = 0; s < h.nums; ++s) h.pcside[fill[h.mapsp1[s]]++] = s;
}

}  // namespace pn

namespace {

template <class T> T* dalloc(pennant_problem* pb, size_t n) {
    void* p = nullptr;
    const size_t bytes = n * sizeof(T) > 0 ? n * sizeof(T) : sizeof(T);
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    if (hipMemset(p, 0, bytes) != hipSuccess) { (void)hipFree(p); return nullptr; }
    pb->allocs.push_back(p);
    return static_cast<T*>(p);
}

template <class T> T* upload(pennant_problem* pb, const std::vector<T>& v) {
    T* d = dalloc<T>(pb, v.size());
    if (d && !v.empty() && hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return nullptr;
    return d;
}

// Reference field catalogue (src/Mesh.cc:262-423): name, entity, components.
struct FieldSpec { const char* name; int kind; int ncomp; };
const FieldSpec kFields[] = {
    {"px", 0, 2}, {"pxp", 0, 2}, {"px0", 0, 2}, {"pu", 0, 2}, {"pu0", 0, 2}, {"pf", 0, 2},
    {"pap", 0, 2}, {"pmaswt", 0, 1},
    {"zx", 1, 2}, {"zxp", 1, 2}, {"zuc", 1, 2}, {"zarea", 1, 1}, {"zvol", 1, 1},
    {"zareap", 1, 1}, {"zvolp", 1, 1}, {"zvol0", 1, 1}, {"zdl", 1, 1}, {"zm", 1, 1},
    {"zr", 1, 1}, {"zrp", 1, 1}, {"ze", 1, 1}, {"zetot", 1, 1}, {"zw", 1, 1},
    {"zwrate", 1, 1}, {"zp", 1, 1}, {"zss", 1, 1}, {"zdu", 1, 1}, {"ztmp", 1, 1},
    {"ex", 2, 2}, {"exp", 2, 2}, {"ssurfp", 2, 2}, {"sfp", 2, 2}, {"sfq", 2, 2}, {"sft", 2, 2},
    {"cqe1", 2, 2}, {"cqe2", 2, 2}, {"sarea", 2, 1}, {"svol", 2, 1}, {"sareap", 2, 1},
    {"svolp", 2, 1}, {"elen", 2, 1}, {"smf", 2, 1}, {"carea", 2, 1}, {"cevol", 2, 1},
    {"cdu", 2, 1}, {"cdiv", 2, 1}, {"ccos", 2, 1}, {"crmu", 2, 1}, {"cw", 2, 1},
};

int64_t count_of(const pennant_problem* pb, int kind) {
    return kind == 0 ? pb->dm.nump : (kind == 1 ? pb->dm.numz : pb->dm.nums);
}

#define CK(x) do { int _e = (int)(x); if (_e) return _e; } while (0)

// One doCycle step through the task entry points; steps as in the header.
int task_step(pennant_problem* pb, int st) {
    const Params& q = pb->prm;
    const pennant_mesh* m = &pb->dm;
    hipStream_t s = pb->stream;
    const double* dt = &pb->ctl->dt;
    switch (st) {
    case 0:  // copy ZVOL -> ZVOL0, src/Hydro.cc:361-368
        return launch_copy_zvol0(pb);
    case 1:  // copies, fills and advPosHalf, src/Hydro.cc:389-437
        CK(launch_point_begin(pb));
        return pennant_hip_adv_pos_half(m, pb->cv("px0"), pb->cv("pu0"), pb->vv("pxp"), dt, s);
    case 2: return pennant_hip_calc_ctrs(m, pb->cv("pxp"), pb->vv("exp"), pb->vv("zxp"), s);
    case 3:
        return pennant_hip_calc_vols(m, pb->cv("pxp"), pb->cv("zxp"), pb->f("sareap"),
                                     pb->f("svolp"), pb->f("zareap"), pb->f("zvolp"),
                                     &pb->ctl->numsbad_pred, pb->ws, s);
    case 4: return pennant_hip_calc_surf_vecs(m, pb->cv("exp"), pb->cv("zxp"), pb->vv("ssurfp"), s);
    case 5: return pennant_hip_calc_edge_len(m, pb->cv("pxp"), pb->f("elen"), s);
    case 6: return pennant_hip_calc_char_len(m, pb->f("sareap"), pb->f("elen"), pb->f("zdl"), s);
    case 7: return pennant_hip_calc_rho(m, pb->f("zm"), pb->f("zvolp"), pb->f("zrp"), s);
    case 8:
        return pennant_hip_calc_crnr_mass(m, pb->f("smf"), pb->f("zrp"), pb->f("zareap"),
                                          pb->f("pmaswt"), s);
    case 9:
        return pennant_hip_calc_state_half(m, pb->f("zr"), pb->f("zvolp"), pb->f("zvol0"),
                                           pb->f("ze"), pb->f("zwrate"), pb->f("zm"),
                                           pb->f("zp"), pb->f("zss"), q.gamma, q.ssmin, dt, s);
    case 10: return pennant_hip_calc_force_pgas(m, pb->cv("ssurfp"), pb->f("zp"), pb->vv("sfp"), s);
    case 11:
        return pennant_hip_calc_force_tts(m, pb->f("sareap"), pb->f("smf"), pb->cv("ssurfp"),
                                          pb->f("zareap"), pb->f("zrp"), pb->f("zss"),
                                          pb->vv("sft"), q.alfa, q.ssmin, s);
    case 12:
        return pennant_hip_set_corner_div(m, pb->cv("exp"), pb->f("elen"), pb->cv("zxp"),
                                          pb->cv("pxp"), pb->cv("pu0"), pb->vv("zuc"),
                                          pb->f("carea"), pb->f("ccos"), pb->f("cdiv"),
                                          pb->f("cevol"), pb->f("cdu"), s);
    case 13:
        return pennant_hip_set_qcn_force(m, pb->f("elen"), pb->f("cdiv"), pb->f("cdu"),
                                         pb->f("cevol"), pb->f("zrp"), pb->f("zss"),
                                         pb->cv("pu0"), pb->f("crmu"), pb->vv("cqe1"),
                                         pb->vv("cqe2"), q.qgamma, q.q1, q.q2, s);
    case 14:
        return pennant_hip_set_force_qcs(m, pb->f("carea"), pb->cv("cqe1"), pb->cv("cqe2"),
                                         pb->f("elen"), pb->f("ccos"), pb->f("cw"),
                                         pb->vv("sfq"), s);
    case 15:
        return pennant_hip_set_vel_diff(m, pb->f("elen"), pb->f("zss"), pb->cv("pxp"),
                                        pb->cv("pu0"), pb->f("ztmp"), pb->f("zdu"), q.q1, q.q2, s);
    case 16:
        return pennant_hip_sum_crnr_force(m, pb->cv("sfp"), pb->cv("sfq"), pb->cv("sft"),
                                          pb->vv("pf"), s);
    case 17:  // one launch per BC list, bcx lists then bcy lists (src/Hydro.cc:773-798)
        for (const BCList& b : pb->bcs)
            CK(pennant_hip_apply_fixed_bc(b.d_pts, b.numb, b.vx, b.vy, pb->vv("pf"),
                                          pb->vv("pu0"), s));
        return 0;
    case 18: return 0;  // bad-side check: done on the device by the end-of-cycle kernel
    case 19:
        CK(pennant_hip_calc_accel(m, pb->cv("pf"), pb->f("pmaswt"), pb->vv("pap"), s));
        return pennant_hip_adv_pos_full(m, pb->cv("px0"), pb->cv("pu0"), pb->cv("pap"),
                                        pb->vv("px"), pb->vv("pu"), dt, s);
    case 20: return pennant_hip_calc_ctrs(m, pb->cv("px"), pb->vv("ex"), pb->vv("zx"), s);
    case 21:
        return pennant_hip_calc_vols(m, pb->cv("px"), pb->cv("zx"), pb->f("sarea"), pb->f("svol"),
                                     pb->f("zarea"), pb->f("zvol"), &pb->ctl->numsbad_corr,
                                     pb->ws, s);
    case 22:
        return pennant_hip_calc_work(m, pb->cv("sfp"), pb->cv("sfq"), pb->cv("pu"), pb->cv("pu0"),
                                     pb->cv("pxp"), pb->f("zw"), pb->f("zetot"), dt, s);
    case 23:
        return pennant_hip_calc_work_rate(m, pb->f("zvol0"), pb->f("zvol"), pb->f("zw"),
                                          pb->f("zp"), pb->f("zwrate"), dt, s);
    case 24: return pennant_hip_calc_energy(m, pb->f("zetot"), pb->f("zm"), pb->f("ze"), s);
    case 25: return pennant_hip_calc_rho(m, pb->f("zm"), pb->f("zvol"), pb->f("zr"), s);
    case 26:
        return pennant_hip_calc_dt_new(m, pb->f("zdl"), pb->f("zdu"), pb->f("zss"), q.cfl,
                                       &pb->ctl->dtnew, pb->ws, s);
    case 27:
        return pennant_hip_calc_dvol(m, pb->f("zvol"), pb->f("zvol0"), &pb->ctl->dvovmax, pb->ws, s);
    case 28:
        return pennant_hip_calc_dt(&pb->ctl->dt, &pb->ctl->dtnew, &pb->ctl->dvovmax, q.cflv,
                                   &pb->ctl->dtrec, s);
    case 29: return 0;
    default: return PENNANT_EINVAL;
    }
}

int copy_ctl_to_host(pennant_problem* pb) {
    CK(hipMemcpyAsync(pb->ctl_host, pb->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, pb->stream));
    return (int)hipStreamSynchronize(pb->stream);
}

int bad_sides_error(pennant_problem* pb) {
    const Ctl& c = *pb->ctl_host;
    if (c.err_cycle < 0) return 0;
    char b[160];
    snprintf(b, sizeof b, "Error: %d negative side volumes on cycle %d",
             c.err_pred ? c.numsbad_pred : c.numsbad_corr, c.err_cycle);
    pb->err = b;
    return PENNANT_EBADSIDES;
}

}  // namespace

static void drop_graph(pennant_problem* pb);

// ===========================================================================
extern "C" {

pennant_problem* pennant_create(const char* deck_text, char* errbuf, int errlen) {
    Params q;
    std::string err;
    pennant_problem* pb = nullptr;
    if (read_params(deck_text, q, err)) {
        HostMesh h;
        build_mesh(q, h);
        pb = pn::create_problem(q, std::move(h), -1, err);
    }
    if (!pb && errbuf && errlen > 0) snprintf(errbuf, errlen, "%s", err.c_str());
    return pb;
}

}  // extern "C"

// Sets up a problem on mesh h: device maps and fields, Mesh::init's geometry,
// side fractions, BC lists, initial state, first dt and the fused-path
// buffers.  npown >= 0 limits the point pass to points [0, npown) (a piece's
// own points; its ghost points come from their owners).
pennant_problem* pn::create_problem(const Params& prm, HostMesh&& hm, int npown, std::string& err) {
    pennant_problem* pb = new pennant_problem();
    auto fail = [&](const std::string& msg) -> pennant_problem* {
        err = msg;
        pennant_destroy(pb);
        return nullptr;
    };
    pb->prm = prm;
    pb->hm = std::move(hm);
    pb->npown = npown;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("no HIP device available");
    const HostMesh& h = pb->hm;
    if (h.maxznump > PENNANT_MAX_ZSIDES) return fail("zones with more than 8 sides are not supported");
    if (hipStreamCreate(&pb->stream) != hipSuccess) return fail("hipStreamCreate failed");

    pennant_mesh& m = pb->dm;
    m.nump = h.nump; m.numz = h.numz; m.nums = h.nums; m.maxznump = h.maxznump;
    m.mapsp1 = upload(pb, h.mapsp1); m.mapsp2 = upload(pb, h.mapsp2); m.mapsz = upload(pb, h.mapsz);
    m.mapss3 = upload(pb, h.mapss3); m.mapss4 = upload(pb, h.mapss4); m.znump = upload(pb, h.znump);
    m.zsbeg = upload(pb, h.zsbeg); m.pcbeg = upload(pb, h.pcbeg); m.pcside = upload(pb, h.pcside);
    if (!m.mapsp1 || !m.pcside) return fail("device allocation failed");
    for (const FieldSpec& fs : kFields) {
        DevField d;
        d.kind = fs.kind; d.ncomp = fs.ncomp; d.c[0] = d.c[1] = nullptr;
        const int64_t n = count_of(pb, fs.kind);
        for (int c = 0; c < fs.ncomp; ++c) {
            d.c[c] = dalloc<double>(pb, (size_t)n);
            if (!d.c[c]) return fail("device allocation failed");
        }
        pb->fields[fs.name] = d;
    }
    pb->ws = dalloc<double>(pb, pennant_hip_reduce_workspace_bytes() / sizeof(double));
    pb->ctl = dalloc<Ctl>(pb, 1);
    if (hipHostMalloc((void**)&pb->ctl_host, sizeof(Ctl), hipHostMallocDefault) != hipSuccess)
        return fail("hipHostMalloc failed");
    std::memset(pb->ctl_host, 0, sizeof(Ctl));
    pb->ctl_host->err_cycle = -1;
    pb->ctl_host->zre_valid = 1;
    if (hipMemcpy(pb->ctl, pb->ctl_host, sizeof(Ctl), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pb->f("px", 0), h.px.data(), h.px.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pb->f("px", 1), h.py.data(), h.py.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return fail("initial upload failed");

    // Mesh::init geometry on PX, src/Mesh.cc:616-623
    hipStream_t s = pb->stream;
    int rc = pennant_hip_calc_ctrs(&m, pb->cv("px"), pb->vv("ex"), pb->vv("zx"), s);
    if (!rc) rc = pennant_hip_calc_vols(&m, pb->cv("px"), pb->cv("zx"), pb->f("sarea"), pb->f("svol"),
                                        pb->f("zarea"), pb->f("zvol"), &pb->ctl->numsbad_corr, pb->ws, s);
    if (!rc) rc = copy_ctl_to_host(pb);
    if (rc) return fail("geometry initialisation failed: " + std::string(hipGetErrorString((hipError_t)rc)));
    if (pb->ctl_host->numsbad_corr > 0) {
        char b[128];
        snprintf(b, sizeof b, "Error: %d negative side volumes during initialization", pb->ctl_host->numsbad_corr);
        return fail(b);
    }
    pb->ctl_host->numsbad_corr = 0;
    if (hipMemcpy(pb->ctl, pb->ctl_host, sizeof(Ctl), hipMemcpyHostToDevice) != hipSuccess)
        return fail("control block upload failed");
    rc = pennant_hip_calc_side_fracs(&m, pb->f("sarea"), pb->f("zarea"), pb->f("smf"), s);

    // HydroBC point lists, src/HydroBC.cc:63-144 (|coordinate - bound| < 1e-12,
    // ascending point order)
    const double eps = 1.e-12;
    auto add_bc = [&](double b, bool xplane) {
        std::vector<int> pts;
        for (int p = 0; p < h.nump; ++p) {
            const double c = xplane ? h.px[p] : h.py[p];
            if (std::fabs(c - b) < eps) pts.push_back(p);
        }
        BCList L;
        L.vx = xplane ? 1. : 0.; L.vy = xplane ? 0. : 1.;
        L.numb = (int)pts.size();
        L.d_pts = upload(pb, pts);
        pb->bcs.push_back(L);
    };
    for (double b : pb->prm.bcx) add_bc(b, true);
    for (double b : pb->prm.bcy) add_bc(b, false);

    if (!rc) rc = launch_init_state(pb);
    if (!rc) rc = launch_first_dt(pb);
    if (!rc) rc = fused_setup(pb);
    if (!rc) rc = (int)hipStreamSynchronize(s);
    if (rc) return fail(std::string("initialisation failed: ") + hipGetErrorString((hipError_t)rc));
    return pb;
}

extern "C" {

void pennant_destroy(pennant_problem* pb) {
    if (!pb) return;
    if (pb->stream) (void)hipStreamSynchronize(pb->stream);
    drop_graph(pb);
    for (void* p : pb->allocs) (void)hipFree(p);
    if (pb->ctl_host) (void)hipHostFree(pb->ctl_host);
    if (pb->stream) (void)hipStreamDestroy(pb->stream);
    delete pb;
}

const char* pennant_error(pennant_problem* pb) { return pb->err.c_str(); }

void pennant_sizes(pennant_problem* pb, int64_t* out4) {
    out4[0] = pb->dm.nump; out4[1] = pb->dm.numz; out4[2] = pb->dm.nums; out4[3] = (int64_t)pb->bcs.size();
}

#define PN_GRAPH_CYCLES 16
#define PN_GRAPH_TASKS 32   // graph_form of captured task-mode cycles

static void drop_graph(pennant_problem* pb) {
    if (pb->gexec) (void)hipGraphExecDestroy(pb->gexec);
    pb->gexec = nullptr;
    pb->graph_form = -1;
}

// The pipelined mixed cycle's look-ahead predictor results no longer match
// the state (the state, the closure, the capture or the launch shapes
// changed, or cycles of another kind ran).
static void drop_lookahead(pennant_problem* pb) {
    // a pipelined launch leaves the look-ahead predictor's zone list behind
    if (pb->la_form >= 0) (void)reset_zone_lists(pb);
    pb->la_form = -1;
}

// Whether cycles of `form` run pipelined: the mixed forms, unless switched
// off or while training tuples are captured.
static bool pipelined(const pennant_problem* pb, int form) {
    return form != FORM_FP64 && pb->tune.pipeline && !pb->cap_buf;
}

// One cycle of `form` (ev as for fused_cycle / fused_cycle_pipe).  A
// pipelined cycle first computes the look-ahead predictor when it is not
// current; any other cycle makes it stale.
static int one_cycle(pennant_problem* pb, hipEvent_t* ev, int form) {
    if (!pipelined(pb, form)) {
        drop_lookahead(pb);
        return fused_cycle(pb, ev, form);
    }
    return fused_cycle_pipe(pb, ev, form);
}

static int ensure_lookahead(pennant_problem* pb, int form) {
    if (!pipelined(pb, form)) {
        drop_lookahead(pb);
        return 0;
    }
    if (pb->la_form == form) return 0;
    CK(fused_prologue(pb, form));
    pb->la_form = form;
    return 0;
}

// Captures PN_GRAPH_CYCLES fused cycles of `form` as one HIP graph (kernel
// nodes only: the fused cycle issues no memset and no host work).
static int build_graph(pennant_problem* pb, int form) {
    drop_graph(pb);
    hipGraph_t g = nullptr;
    const bool pipe = pipelined(pb, form);
    CK(hipStreamBeginCapture(pb->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int k = 0; k < PN_GRAPH_CYCLES && !rc; ++k)
        rc = pipe ? fused_cycle_pipe(pb, nullptr, form) : fused_cycle(pb, nullptr, form);
    const int rc2 = (int)hipStreamEndCapture(pb->stream, &g);
    if (rc || rc2) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : rc2;
    }
    rc = (int)hipGraphInstantiate(&pb->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (rc) { pb->gexec = nullptr; return rc; }
    pb->graph_form = form + (pipe ? 16 : 0);
    return 0;
}

// One task-mode cycle: the 30 doCycle steps through the task entry points,
// then the end-of-cycle update (time, cycle, stop test, next dt).
static int task_cycle(pennant_problem* pb) {
    for (int st = 0; st < 30; ++st) CK(task_step(pb, st));
    return launch_end_cycle(pb, 0);
}

// Whether `n` task-mode cycles from the state in ctl_host all run before the
// stop condition.  cstop is exact.  For tstop: the next cycle's dt is known,
// and each later dt is at most dtfac times the one before
// (calcGlobalDt, src/Driver.cc:225-231) and at most dtmax, so the time after
// n cycles is bounded; a margin covers the rounding of the time sums.
static bool task_run_is_safe(const pennant_problem* pb, int n) {
    const Ctl& c = *pb->ctl_host;
    const Params& q = pb->prm;
    if (c.done || (long long)c.cycle + n > (long long)q.cstop) return false;
    const double f = q.dtfac > 1. ? q.dtfac : 1.;
    double dt = c.dt, span = 0.;
    for (int k = 0; k < n; ++k) {
        span += dt;
        dt = std::fmin(dt * f, q.dtmax);
    }
    if (!(span >= 0.) || !std::isfinite(span)) return false;
    return c.time + span * (1. + 1.e-6) + 1.e-12 * std::fabs(q.tstop) < q.tstop;
}

// Captures PN_GRAPH_CYCLES task-mode cycles as one HIP graph (kernel nodes
// only: the copies and fills of doCycle are kernels too).
static int build_task_graph(pennant_problem* pb) {
    drop_graph(pb);
    hipGraph_t g = nullptr;
    CK(hipStreamBeginCapture(pb->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int k = 0; k < PN_GRAPH_CYCLES && !rc; ++k) rc = task_cycle(pb);
    const int rc2 = (int)hipStreamEndCapture(pb->stream, &g);
    if (rc || rc2) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : rc2;
    }
    rc = (int)hipGraphInstantiate(&pb->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (rc) { pb->gexec = nullptr; return rc; }
    pb->graph_form = PN_GRAPH_TASKS;
    return 0;
}

// fused-cycle form of a driver mode, or -1 for task mode / unknown modes
static int fused_form(int mode) {
    switch (mode) {
        case PENNANT_MODE_FUSED: return FORM_FP64;
        case PENNANT_MODE_MIXED: return FORM_MIXED;
        case PENNANT_MODE_LEARNED: return FORM_LEARNED;
        default: return -1;
    }
}

int pennant_launch_cycles(pennant_problem* pb, int ncycles, int mode) {
    const int form = fused_form(mode);
    if (pb->smf_stale) CK(refresh_side_fracs(pb));
    // the FP64 and task forms read zr and ze; mixed cycles leave them derived
    if (form == FORM_FP64 || mode == PENNANT_MODE_TASKS) CK(materialize_zre(pb));
    else if (form > 0) pb->zre_stale = true;
    if (form >= 0) {
        if (form == FORM_LEARNED && !pb->closure) return PENNANT_EINVAL;
        int left = ncycles;
        if (pb->tune.graph && left >= PN_GRAPH_CYCLES) {
            const int key = form + (pipelined(pb, form) ? 16 : 0);
            if (pb->graph_form != key || !pb->gexec) CK(build_graph(pb, form));
            CK(ensure_lookahead(pb, form));
            for (; left >= PN_GRAPH_CYCLES; left -= PN_GRAPH_CYCLES) CK(hipGraphLaunch(pb->gexec, pb->stream));
        }
        if (left > 0) CK(ensure_lookahead(pb, form));
        for (; left > 0; --left) CK(one_cycle(pb, nullptr, form));
        return 0;
    }
    if (mode != PENNANT_MODE_TASKS) return PENNANT_EINVAL;
    drop_lookahead(pb);
    // Task mode mirrors Driver::run, one launch per task.  The task entry
    // points have no stop predicate, so the host reads the control block
    // before every run of cycles it queues: a run of PN_GRAPH_CYCLES cycles
    // replays one captured HIP graph when no cycle of it can reach the stop
    // condition (task_run_is_safe), otherwise cycles go one at a time.
    for (int k = 0; k < ncycles;) {
        CK(copy_ctl_to_host(pb));
        if (pb->ctl_host->done) break;
        if (pb->tune.graph && ncycles - k >= PN_GRAPH_CYCLES &&
            task_run_is_safe(pb, PN_GRAPH_CYCLES)) {
            if (pb->graph_form != PN_GRAPH_TASKS || !pb->gexec) CK(build_task_graph(pb));
            CK(hipGraphLaunch(pb->gexec, pb->stream));
            k += PN_GRAPH_CYCLES;
            continue;
        }
        CK(task_cycle(pb));
        ++k;
    }
    return 0;
}

int pennant_time_cycles(pennant_problem* pb, int ncycles, int mode, double* total_ms,
                        double* phase_ms) {
    if (ncycles < 1) return PENNANT_EINVAL;
    const int form = fused_form(mode);
    const bool phases = (form >= 0 && phase_ms);
    // per cycle: events before K1, the learned pass, K2, K3, K4 and after K4
    // (the pipelined cycle records the first five)
    const size_t E = 6;
    std::vector<hipEvent_t> ev(phases ? E * (size_t)ncycles + 2 : 2);
    for (hipEvent_t& e : ev) CK(hipEventCreate(&e));
    int rc = pb->smf_stale ? refresh_side_fracs(pb) : 0;
    if (!rc && phases) rc = ensure_lookahead(pb, form);
    if (!rc && phases && form == FORM_FP64) rc = materialize_zre(pb);
    if (phases && form > 0) pb->zre_stale = true;
    const bool pipe = phases && pipelined(pb, form);
    if (!rc) rc = (int)hipEventRecord(ev[0], pb->stream);
    if (!rc) {
        if (phases) {
            for (int k = 0; k < ncycles && !rc; ++k) rc = one_cycle(pb, &ev[2 + E * (size_t)k], form);
        } else {
            rc = pennant_launch_cycles(pb, ncycles, mode);
        }
    }
    if (!rc) rc = (int)hipEventRecord(ev[1], pb->stream);
    if (!rc) rc = (int)hipEventSynchronize(ev[1]);
    float ms = 0.f;
    if (!rc) rc = (int)hipEventElapsedTime(&ms, ev[0], ev[1]);
    if (!rc) *total_ms = ms;
    if (!rc && phases) {
        // reported as K1 (or the pipelined zone pass), K2, K3, K4 (with the
        // pipelined redo) and the learned pass.  Unpipelined intervals: K1,
        // learned pass, K2, K3, K4; pipelined: K2, zone pass, K4 and redo,
        // learned pass
        static const int kOrder[5] = {0, 4, 1, 2, 3}, kPipe[4] = {1, 0, 3, 4};
        const int* slot = pipe ? kPipe : kOrder;
        const int nint = pipe ? 4 : 5;
        for (int j = 0; j < 5; ++j) phase_ms[j] = 0.;
        for (int k = 0; k < ncycles && !rc; ++k)
            for (int j = 0; j < nint && !rc; ++j) {
                // outside the learned form the learned-pass interval holds no
                // kernel (only the events' own overhead): it reports 0
                if (slot[j] == 4 && form != FORM_LEARNED) continue;
                rc = (int)hipEventElapsedTime(&ms, ev[2 + E * (size_t)k + j], ev[2 + E * (size_t)k + j + 1]);
                phase_ms[slot[j]] += ms;
            }
    }
    for (hipEvent_t& e : ev) (void)hipEventDestroy(e);
    if (rc) return rc;
    return pennant_sync(pb);
}

int pennant_set_tuning(pennant_problem* pb, const char* key, int value) {
    drop_graph(pb);
    drop_lookahead(pb);
    if (!strcmp(key, "graph")) {
        if (value != 0 && value != 1) return PENNANT_EINVAL;
        pb->tune.graph = value;
        return 0;
    }
    if (!strcmp(key, "pipeline")) {
        if (value != 0 && value != 1) return PENNANT_EINVAL;
        pb->tune.pipeline = value;
        return 0;
    }
    if (!strcmp(key, "k1_layout")) {
        if (value != 0 && value != 1) return PENNANT_EINVAL;
        pb->tune.k1_layout = value;
        return 0;
    }
    if (!strcmp(key, "max_blocks")) {
        if (value < 1 || value > PN_MAX_GRID) return PENNANT_EINVAL;
        pb->tune.max_blocks = value;
        return 0;
    }
    if (!strcmp(key, "xcd")) {
        if (value != 0 && value != 1) return PENNANT_EINVAL;
        pb->tune.xcd = value;
        return 0;
    }
    if (!strcmp(key, "rect_maps")) {
        if (value != 0 && value != 1) return PENNANT_EINVAL;
        pb->tune.rect_maps = value;
        return 0;
    }
    return PENNANT_EINVAL;
}

int pennant_get_tuning(pennant_problem* pb, const char* key) {
    if (!strcmp(key, "k1_layout")) return pb->tune.k1_layout;
    if (!strcmp(key, "max_blocks")) return pb->tune.max_blocks;
    if (!strcmp(key, "xcd")) return pb->tune.xcd;
    if (!strcmp(key, "graph")) return pb->tune.graph;
    if (!strcmp(key, "pipeline")) return pb->tune.pipeline;
    // whether the fused kernels compute the zones' points and corner slots
    if (!strcmp(key, "rect_maps")) return (pb->tune.rect_maps && pb->rect_nzx > 0) ? 1 : 0;
    return PENNANT_EINVAL;
}

int pennant_sync(pennant_problem* pb) {
    CK(copy_ctl_to_host(pb));
    return bad_sides_error(pb);
}

// Times `cycles` cycles (after one untimed cycle) at every launch shape from
// the same snapshot of the state, keeps the fastest shape and restores the
// snapshot, so the problem does not advance.  Results do not depend on the
// shape, only the time does.
int pennant_autotune(pennant_problem* pb, int mode, int cycles, double* best_ms) {
    const int form = fused_form(mode);
    if (form < 0 || cycles < 1) return PENNANT_EINVAL;
    if (form == FORM_LEARNED && !pb->closure) return PENNANT_EINVAL;
    const size_t Z = (size_t)pb->dm.numz * 8, P = (size_t)pb->dm.nump * 8;
    const std::pair<void*, size_t> st[] = {
        {pb->f("px", 0), P}, {pb->f("px", 1), P}, {pb->f("pu", 0), P}, {pb->f("pu", 1), P},
        {pb->f("zr"), Z}, {pb->f("ze"), Z}, {pb->f("zetot"), Z}, {pb->f("zwrate"), Z},
        {pb->f("zvol"), Z}, {pb->f("zp"), Z}, {pb->ctl, sizeof(Ctl)},
        // the closure and capture counters, so the timing cycles do not count
        {pb->cl_counters, 5 * sizeof(unsigned long long)}};
    std::vector<void*> copy;
    int rc = 0;
    for (const auto& a : st) {
        void* c = nullptr;
        if (!rc) rc = (int)hipMalloc(&c, a.second);
        if (!rc) copy.push_back(c);
        if (!rc) rc = (int)hipMemcpyAsync(c, a.first, a.second, hipMemcpyDeviceToDevice, pb->stream);
    }
    auto restore = [&]() {
        drop_lookahead(pb);
        pb->zre_stale = true;        // the restored flag decides; materialising checks it
        int r = reset_zone_lists(pb);
        for (size_t i = 0; i < copy.size() && !r; ++i)
            r = (int)hipMemcpyAsync(st[i].first, copy[i], st[i].second, hipMemcpyDeviceToDevice, pb->stream);
        if (!r) r = (int)hipStreamSynchronize(pb->stream);
        return r;
    };
    const Tuning keep = pb->tune;
    Tuning best = keep;
    double bms = 1e300;
    drop_graph(pb);
    const int grids[] = {256, 512, 1024, 2048, 4096};
    // the mixed forms also try the pipelined cycle (it pays where the
    // volume-change limit seldom sets dt)
    const int npipe = (form == FORM_FP64 || pb->cap_buf) ? 1 : 2;
    for (int layout = 0; layout <= (form == FORM_FP64 ? 1 : 0) && !rc; ++layout)
      for (int pipe = 0; pipe < npipe && !rc; ++pipe)
        for (int xcd = 0; xcd <= 1 && !rc; ++xcd)
            for (int g : grids) {
                if (rc) break;
                pb->tune.k1_layout = layout;
                pb->tune.pipeline = (npipe == 2) ? pipe : keep.pipeline;
                pb->tune.xcd = xcd;
                pb->tune.max_blocks = g;
                pb->tune.graph = 0;               // direct launches while shapes change
                rc = restore();
                if (!rc) rc = pennant_launch_cycles(pb, 1, mode);
                double ms = 0.;
                if (!rc) rc = pennant_time_cycles(pb, cycles, mode, &ms, nullptr);
                if (rc == PENNANT_EBADSIDES) rc = 0;
                if (!rc && ms < bms) { bms = ms; best = pb->tune; }
            }
    pb->tune = (bms < 1e300) ? best : keep;
    pb->tune.graph = keep.graph;
    drop_graph(pb);
    const int rc2 = restore();
    for (void* c : copy) (void)hipFree(c);
    if (best_ms) *best_ms = bms;
    return rc ? rc : rc2;
}

// HIP errors (positive codes) as negative return values, for the entry
// points whose non-negative results are counts
#define NCK(x) do { int _e = (int)(x); if (_e) return _e > 0 ? -_e : _e; } while (0)

int pennant_run(pennant_problem* pb, int ncycles, int mode) {
    NCK(copy_ctl_to_host(pb));
    const int c0 = pb->ctl_host->cycle;
    int left = ncycles, chunk = 32;
    while (left > 0 && !pb->ctl_host->done) {
        const int k = left < chunk ? left : chunk;
        NCK(pennant_launch_cycles(pb, k, mode));
        NCK(pennant_sync(pb));
        left -= k;
        chunk = chunk < 1024 ? 2 * chunk : 1024;
    }
    return pb->ctl_host->cycle - c0;
}

// Driver::run, src/Driver.cc:102-187: the cycle loop with the report line
// after loop cycle 0 and after every cycle whose count is a multiple of
// dtreport (src/Driver.cc:147-156, printed by reportMeasurementTask,
// src/Driver.cc:330-344), then the end-of-run summary (src/Driver.cc:169-180).
// The host never waits per cycle: the cycles up to each report boundary are
// queued, followed by an event and an asynchronous copy of the control block
// into a pinned slot, with up to PN_DRIVE_WINDOW boundaries in flight.  The
// device predicate makes cycles queued past the stop condition no-ops, as
// the reference's predicated launches are.
#define PN_DRIVE_WINDOW 32
#define PN_DRIVE_AHEAD 2048   // most cycles queued ahead of the last report read

namespace {
struct DriveSeg {
    int last;          // loop cycle index of the boundary
    hipEvent_t end;    // recorded after the boundary's cycle
    hipEvent_t copied; // recorded after the control block copy
};

void emit(pennant_report_fn report, void* user, const char* text) {
    if (report) {
        report(text, user);
    } else {
        std::fputs(text, stdout);
        std::fflush(stdout);
    }
}
}  // namespace

int pennant_drive(pennant_problem* pb, int mode, pennant_report_fn report, void* user,
                  double* run_us) {
    if (mode < PENNANT_MODE_TASKS || mode > PENNANT_MODE_LEARNED) return PENNANT_EINVAL;
    if (mode == PENNANT_MODE_LEARNED && !pb->closure) return PENNANT_EINVAL;
    const Params& q = pb->prm;
    const int dtreport = q.dtreport > 0 ? q.dtreport : 1;
    CK(copy_ctl_to_host(pb));
    const int c0 = pb->ctl_host->cycle;
    bool done = pb->ctl_host->done != 0;
    Ctl* slots = nullptr;
    CK(hipHostMalloc((void**)&slots, sizeof(Ctl) * PN_DRIVE_WINDOW, hipHostMallocDefault));
    std::vector<DriveSeg> ring(PN_DRIVE_WINDOW);
    int rc = 0;
    int ncreated = 0;
    for (DriveSeg& d : ring) {
        if (!rc) rc = (int)hipEventCreate(&d.end);
        if (!rc) rc = (int)hipEventCreate(&d.copied);
        if (!rc) ++ncreated;
    }
    hipEvent_t start = nullptr;
    if (!rc) rc = (int)hipEventCreate(&start);
    if (!rc) rc = (int)hipEventRecord(start, pb->stream);
    // Times are kept as milliseconds since `start`, which is never re-recorded
    // (the ring's events are, once their slot comes round again).
    float prev_ms = 0.f;            // the last measurement (src/Driver.cc:122-123)
    float stop_ms = 0.f;            // end of the last cycle that ran
    int next = c0;                  // next loop cycle index to queue
    int seen = c0;                  // cycle count of the last snapshot read
    int head = 0, tail = 0;         // ring of queued boundaries [head, tail)
    char line[160];
    while (!rc && (head < tail || (!done && next < q.cstop))) {
        // queue boundaries while the window allows
        while (!rc && !done && next < q.cstop && tail - head < PN_DRIVE_WINDOW &&
               (head == tail || next - ring[head % PN_DRIVE_WINDOW].last <= PN_DRIVE_AHEAD)) {
            // the report boundary at or after loop cycle `next`
            int last = (next == 0) ? 0 : ((next + 1 + dtreport - 1) / dtreport) * dtreport - 1;
            if (last < next) last = next;
            if (last > q.cstop - 1) last = q.cstop - 1;
            DriveSeg& d = ring[tail % PN_DRIVE_WINDOW];
            rc = pennant_launch_cycles(pb, last - next + 1, mode);
            if (!rc) rc = (int)hipEventRecord(d.end, pb->stream);
            if (!rc) rc = (int)hipMemcpyAsync(&slots[tail % PN_DRIVE_WINDOW], pb->ctl, sizeof(Ctl),
                                              hipMemcpyDeviceToHost, pb->stream);
            if (!rc) rc = (int)hipEventRecord(d.copied, pb->stream);
            d.last = last;
            next = last + 1;
            ++tail;
        }
        if (rc || head == tail) break;
        // read the oldest boundary
        DriveSeg& d = ring[head % PN_DRIVE_WINDOW];
        rc = (int)hipEventSynchronize(d.copied);
        if (rc) break;
        const Ctl& c = slots[head % PN_DRIVE_WINDOW];
        const bool report_here = (d.last == 0) || ((d.last + 1) % dtreport == 0);
        float end_ms = 0.f;
        rc = (int)hipEventElapsedTime(&end_ms, start, d.end);
        if (rc) break;
        if (c.cycle > seen) {
            // cycles ran in this segment: the run time extends to its end
            stop_ms = end_ms;
            seen = c.cycle;
        }
        if (c.cycle == d.last + 1 && report_here) {
            // this boundary's cycle ran
            snprintf(line, sizeof line, "End cycle %6d, time = %11.5g, dt = %11.5g, wall = %11lld us\n",
                     c.cycle, c.time, c.dtlast, (long long)((end_ms - prev_ms) * 1e3f));
            emit(report, user, line);
            prev_ms = end_ms;
        }
        if (c.done) done = true;
        ++head;
    }
    const float total_ms = stop_ms;
    if (!rc) rc = copy_ctl_to_host(pb);
    if (!rc) rc = bad_sides_error(pb);
    if (!rc) {
        const Ctl& c = *pb->ctl_host;
        snprintf(line, sizeof line, "\nRun complete\n");
        emit(report, user, line);
        snprintf(line, sizeof line, "cycle = %6d,        cstop = %8d\n", c.cycle, q.cstop);
        emit(report, user, line);
        snprintf(line, sizeof line, "time = %14.6g, tstop = %8.6g\n\n", c.time, q.tstop);
        emit(report, user, line);
        emit(report, user, "************************************\n");
        snprintf(line, sizeof line, "hydro cycle run time= %14.8g us\n", (double)total_ms * 1e3);
        emit(report, user, line);
        emit(report, user, "************************************\n");
        if (run_us) *run_us = (double)total_ms * 1e3;
    }
    (void)hipStreamSynchronize(pb->stream);
    for (int i = 0; i < ncreated; ++i) {
        (void)hipEventDestroy(ring[i].end);
        (void)hipEventDestroy(ring[i].copied);
    }
    if (start) (void)hipEventDestroy(start);
    (void)hipHostFree(slots);
    if (rc) return rc > 0 ? -rc : rc;   // HIP errors as negative codes
    return pb->ctl_host->cycle - c0;
}

int pennant_lookahead_redos(pennant_problem* pb, int64_t* count) {
    CK(copy_ctl_to_host(pb));
    *count = pb->ctl_host->la_redos;
    return 0;
}

int pennant_scalars(pennant_problem* pb, double* o) {
    CK(copy_ctl_to_host(pb));
    const Ctl& c = *pb->ctl_host;
    o[0] = c.time; o[1] = c.dtlast; o[2] = c.dtrec; o[3] = c.cycle; o[4] = c.dtnew;
    o[5] = c.dvovmax; o[6] = c.numsbad_pred; o[7] = c.numsbad_corr; o[8] = c.dt; o[9] = c.done;
    return 0;
}

int pennant_set_scalars(pennant_problem* pb, const double* in) {
    drop_lookahead(pb);
    CK(copy_ctl_to_host(pb));
    Ctl& c = *pb->ctl_host;
    c.time = in[0]; c.dtlast = in[1]; c.dtrec = in[2]; c.cycle = (int)in[3]; c.dtnew = in[4];
    c.dvovmax = in[5]; c.numsbad_pred = (int)in[6]; c.numsbad_corr = (int)in[7]; c.dt = in[8];
    c.done = (int)in[9]; c.err_cycle = -1; c.err_pred = 0;
    CK(hipMemcpyAsync(pb->ctl, pb->ctl_host, sizeof(Ctl), hipMemcpyHostToDevice, pb->stream));
    return (int)hipStreamSynchronize(pb->stream);
}

int pennant_field_info(pennant_problem* pb, const char* name, int64_t* count, int* ncomp) {
    auto it = pb->fields.find(name);
    if (it == pb->fields.end()) return PENNANT_EINVAL;
    *count = count_of(pb, it->second.kind);
    *ncomp = it->second.ncomp;
    return 0;
}

int pennant_get_field(pennant_problem* pb, const char* name, double* out) {
    CK(materialize_zre(pb));
    auto it = pb->fields.find(name);
    if (it == pb->fields.end()) return PENNANT_EINVAL;
    const DevField& d = it->second;
    const int64_t n = count_of(pb, d.kind);
    CK(hipStreamSynchronize(pb->stream));
    std::vector<double> tmp((size_t)n);
    for (int c = 0; c < d.ncomp; ++c) {
        CK(hipMemcpy(tmp.data(), d.c[c], (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i) out[i * d.ncomp + c] = tmp[(size_t)i];
    }
    return 0;
}

int pennant_set_field(pennant_problem* pb, const char* name, const double* in) {
    drop_lookahead(pb);
    CK(materialize_zre(pb));
    auto it = pb->fields.find(name);
    if (it == pb->fields.end()) return PENNANT_EINVAL;
    const DevField& d = it->second;
    const int64_t n = count_of(pb, d.kind);
    CK(hipStreamSynchronize(pb->stream));
    std::vector<double> tmp((size_t)n);
    for (int c = 0; c < d.ncomp; ++c) {
        for (int64_t i = 0; i < n; ++i) tmp[(size_t)i] = in[i * d.ncomp + c];
        CK(hipMemcpy(d.c[c], tmp.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    }
    // the mixed forms read an FP32 copy of smf
    if (!strcmp(name, "smf")) CK(refresh_side_fracs(pb));
    return 0;
}

double* pennant_field_device(pennant_problem* pb, const char* name, int comp) {
    // the caller may write through the pointer
    drop_lookahead(pb);
    if (materialize_zre(pb)) return nullptr;
    auto it = pb->fields.find(name);
    if (it == pb->fields.end() || comp < 0 || comp >= it->second.ncomp) return nullptr;
    if (!strcmp(name, "smf")) pb->smf_stale = true;   // rebuilt before the next fused cycle
    return it->second.c[comp];
}

}  // extern "C"

namespace {
const std::vector<int>* host_map(const HostMesh& h, const char* name) {
    const std::vector<int>* v = nullptr;
    if (!strcmp(name, "mapsp1")) v = &h.mapsp1;
    else if (!strcmp(name, "mapsp2")) v = &h.mapsp2;
    else if (!strcmp(name, "mapsz")) v = &h.mapsz;
    else if (!strcmp(name, "mapss3")) v = &h.mapss3;
    else if (!strcmp(name, "mapss4")) v = &h.mapss4;
    else if (!strcmp(name, "znump")) v = &h.znump;
    else if (!strcmp(name, "zsbeg")) v = &h.zsbeg;
    else if (!strcmp(name, "pcbeg")) v = &h.pcbeg;
    else if (!strcmp(name, "pcside")) v = &h.pcside;
    return v;
}
}  // namespace

struct pennant_host_mesh {
    Params prm;
    HostMesh hm;
};

extern "C" {

int pennant_get_map(pennant_problem* pb, const char* name, int32_t* out, int64_t* count) {
    const std::vector<int>* v = host_map(pb->hm, name);
    if (!v) return PENNANT_EINVAL;
    *count = (int64_t)v->size();
    if (out) std::memcpy(out, v->data(), v->size() * sizeof(int32_t));
    return 0;
}

pennant_host_mesh* pennant_host_mesh_create(const char* deck_text, char* errbuf, int errlen) {
    pennant_host_mesh* hm = new pennant_host_mesh();
    std::string err;
    if (!read_params(deck_text, hm->prm, err)) {
        if (errbuf && errlen > 0) snprintf(errbuf, errlen, "%s", err.c_str());
        delete hm;
        return nullptr;
    }
    build_mesh(hm->prm, hm->hm);
    return hm;
}

int pennant_host_mesh_get(pennant_host_mesh* hm, const char* name, void* out, int64_t* count) {
    const HostMesh& h = hm->hm;
    if (!strcmp(name, "px") || !strcmp(name, "py")) {
        const std::vector<double>& v = (name[1] == 'x') ? h.px : h.py;
        *count = (int64_t)v.size();
        if (out) std::memcpy(out, v.data(), v.size() * sizeof(double));
        return 0;
    }
    const std::vector<int>* v = host_map(h, name);
    if (!v) return PENNANT_EINVAL;
    *count = (int64_t)v->size();
    if (out) std::memcpy(out, v->data(), v->size() * sizeof(int32_t));
    return 0;
}

void pennant_host_mesh_destroy(pennant_host_mesh* hm) { delete hm; }

const pennant_mesh* pennant_mesh_desc(pennant_problem* pb) { return &pb->dm; }
hipStream_t pennant_stream(pennant_problem* pb) { return pb->stream; }

int pennant_begin_cycle(pennant_problem* pb) {
    CK(copy_ctl_to_host(pb));
    return pb->ctl_host->done ? 1 : 0;
}

int pennant_step(pennant_problem* pb, int step) {
    drop_lookahead(pb);
    CK(materialize_zre(pb));
    CK(task_step(pb, step));
    CK(copy_ctl_to_host(pb));
    if (step == 18 && pb->ctl_host->numsbad_pred > 0) return PENNANT_EBADSIDES;
    if (step == 29 && pb->ctl_host->numsbad_corr > 0) return PENNANT_EBADSIDES;
    return 0;
}

int pennant_end_cycle(pennant_problem* pb) {
    drop_lookahead(pb);
    CK(materialize_zre(pb));
    CK(launch_end_cycle(pb, 0));
    return pennant_sync(pb);
}

int pennant_write_xy(pennant_problem* pb, const char* path) {
    const int64_t nz = pb->dm.numz;
    std::vector<double> zr(nz), ze(nz), zp(nz);
    CK(pennant_get_field(pb, "zr", zr.data()));
    CK(pennant_get_field(pb, "ze", ze.data()));
    CK(pennant_get_field(pb, "zp", zp.data()));
    FILE* f = std::fopen(path, "w");
    if (!f) return PENNANT_EINVAL;
    const std::vector<double>* a[3] = {&zr, &ze, &zp};
    const char* hdr[3] = {"#  zr", "#  ze", "#  zp"};
    for (int k = 0; k < 3; ++k) {
        std::fprintf(f, "%s\n", hdr[k]);
        for (int64_t z = 0; z < nz; ++z) std::fprintf(f, "%5lld%18.8e\n", (long long)(z + 1), (*a[k])[z]);
    }
    std::fclose(f);
    return 0;
}

// EnSight Gold export, the file set of src/ExportGold.cc:31-307: <base>.case,
// <base>.geo (coordinates; then triangles, quads and n-sided zones, each group
// listing its zone ids and its corner points, 1-based) and one per-element
// scalar file per zone field.  The reference writes through iostreams in
// scientific notation; the printf conversions below produce the same text
// (setw(w) << int is %wd, setprecision(p) << double is %w.pe).
int pennant_write_gold_host(const char* base, int32_t cycle, double time, const double* zr,
                            const double* ze, const double* zp, const int32_t* znump,
                            int32_t numz, const double* px, const double* py, int32_t nump,
                            const int32_t* mapsp1) {
    if (!base || numz < 0 || nump < 0) return PENNANT_EINVAL;
    const std::string b(base);
    // zones grouped by size (ExportGold::sortZones), with their first side
    std::vector<int> tris, quads, others;
    std::vector<int64_t> zs(numz);
    int64_t sc = 0;
    for (int z = 0; z < numz; ++z) {
        if (znump[z] == 3) tris.push_back(z);
        else if (znump[z] == 4) quads.push_back(z);
        else others.push_back(z);
        zs[z] = sc;
        sc += znump[z];
    }
    FILE* f = std::fopen((b + ".case").c_str(), "w");
    if (!f) return PENNANT_EINVAL;
    std::fprintf(f, "#\n# Created by PENNANT\n#\nFORMAT\ntype: ensight gold\nGEOMETRY\nmodel: %s.geo\n"
                    "VARIABLE\nscalar per element: zr %s.zr\nscalar per element: ze %s.ze\n"
                    "scalar per element: zp %s.zp\n", base, base, base, base);
    std::fclose(f);

    f = std::fopen((b + ".geo").c_str(), "w");
    if (!f) return PENNANT_EINVAL;
    std::fprintf(f, "cycle = %8d\nt = %15.8e\nnode id assign\nelement id given\npart\n%10d\n"
                    "universe\ncoordinates\n%10d\n", cycle, time, 1, nump);
    for (int p = 0; p < nump; ++p) std::fprintf(f, "%12.5e\n", px[p]);
    for (int p = 0; p < nump; ++p) std::fprintf(f, "%12.5e\n", py[p]);
    for (int p = 0; p < nump; ++p) std::fprintf(f, "%12.5e\n", 0.);
    auto group = [&](const char* kind, const std::vector<int>& zl, bool nsided) {
        if (zl.empty()) return;
        std::fprintf(f, "%s\n%10d\n", kind, (int)zl.size());
        for (int z : zl) std::fprintf(f, "%10d\n", z + 1);
        if (nsided)
            for (int z : zl) std::fprintf(f, "%10d\n", znump[z]);
        for (int z : zl) {
            for (int i = 0; i < znump[z]; ++i) std::fprintf(f, "%10d", mapsp1[zs[z] + i] + 1);
            std::fputc('\n', f);
        }
    };
    group("tria3", tris, false);
    group("quad4", quads, false);
    group("nsided", others, true);
    std::fclose(f);

    const char* names[3] = {"zr", "ze", "zp"};
    const double* vars[3] = {zr, ze, zp};
    for (int k = 0; k < 3; ++k) {
        f = std::fopen((b + "." + names[k]).c_str(), "w");
        if (!f) return PENNANT_EINVAL;
        std::fprintf(f, "%s\npart\n%10d\n", names[k], 1);
        const char* kinds[3] = {"tria3", "quad4", "nsided"};
        const std::vector<int>* lists[3] = {&tris, &quads, &others};
        for (int g = 0; g < 3; ++g) {
            if (lists[g]->empty()) continue;
            std::fprintf(f, "%s\n", kinds[g]);
            for (int z : *lists[g]) std::fprintf(f, "%12.5e\n", vars[k][z]);
        }
        std::fclose(f);
    }
    return 0;
}

const void* pennant_mlp_device_pack(pennant_mlp* m);

int pennant_set_closure(pennant_problem* pb, pennant_mlp* m) {
    pb->closure = m ? pennant_mlp_device_pack(m) : nullptr;
    drop_graph(pb);
    drop_lookahead(pb);
    return 0;
}

int pennant_closure_stats(pennant_problem* pb, int64_t* out2) {
    unsigned long long c[3];
    CK(hipMemcpyAsync(c, pb->cl_counters, sizeof c, hipMemcpyDeviceToHost, pb->stream));
    CK(hipStreamSynchronize(pb->stream));
    out2[0] = (int64_t)c[1];
    out2[1] = (int64_t)c[2];
    return 0;
}

int pennant_set_capture(pennant_problem* pb, float* dev_buf, int64_t cap) {
    if (cap < 0 || (dev_buf == nullptr) != (cap == 0)) return PENNANT_EINVAL;
    CK(hipStreamSynchronize(pb->stream));
    pb->cap_buf = dev_buf;
    pb->cap = cap;
    drop_graph(pb);
    drop_lookahead(pb);
    const unsigned long long zero = 0;
    return (int)hipMemcpy(pb->cl_counters, &zero, sizeof zero, hipMemcpyHostToDevice);
}

int pennant_capture_count(pennant_problem* pb, int64_t* count) {
    unsigned long long c = 0;
    CK(hipMemcpyAsync(&c, pb->cl_counters, sizeof c, hipMemcpyDeviceToHost, pb->stream));
    CK(hipStreamSynchronize(pb->stream));
    *count = (int64_t)c;
    return 0;
}

int pennant_energy(pennant_problem* pb, double* internal, double* kinetic) {
    const HostMesh& h = pb->hm;
    const int nz = h.numz, np = h.nump, ns = h.nums;
    std::vector<double> zetot(nz), zm(nz), zvol(nz), px(np), pux(np), puy(np), smf(ns);
    CK(pennant_get_field(pb, "zetot", zetot.data()));
    CK(pennant_get_field(pb, "zm", zm.data()));
    CK(pennant_get_field(pb, "zvol", zvol.data()));
    CK(pennant_get_field(pb, "smf", smf.data()));
    CK(hipMemcpy(px.data(), pb->f("px", 0), sizeof(double) * np, hipMemcpyDeviceToHost));
    CK(hipMemcpy(pux.data(), pb->f("pu", 0), sizeof(double) * np, hipMemcpyDeviceToHost));
    CK(hipMemcpy(puy.data(), pb->f("pu", 1), sizeof(double) * np, hipMemcpyDeviceToHost));
    // planar zone areas of the current positions
    std::vector<double> py(np);
    CK(hipMemcpy(py.data(), pb->f("px", 1), sizeof(double) * np, hipMemcpyDeviceToHost));
    double ei = 0., ek = 0.;
    for (int z = 0; z < nz; ++z) ei += zetot[z];
    for (int z = 0; z < nz; ++z) {
        const int sb = h.zsbeg[z], n = h.znump[z];
        double cx = 0., cy = 0.;
        for (int k = 0; k < n; ++k) {
            cx += px[h.mapsp1[sb + k]] / n;
            cy += py[h.mapsp1[sb + k]] / n;
        }
        double zarea = 0.;
        for (int k = 0; k < n; ++k) {
            const int p1 = h.mapsp1[sb + k], p2 = h.mapsp2[sb + k];
            zarea += 0.5 * ((px[p2] - px[p1]) * (cy - py[p1]) - (py[p2] - py[p1]) * (cx - px[p1]));
        }
        for (int k = 0; k < n; ++k) {
            const int s = sb + k, s3 = h.mapss3[s], p = h.mapsp1[s];
            const double cvol = zarea * px[p] * 0.5 * (smf[s] + smf[s3]);
            ek += zm[z] * cvol / zvol[z] * 0.5 * (pux[p] * pux[p] + puy[p] * puy[p]);
        }
    }
    *internal = ei;
    *kinetic = ek;
    return 0;
}

int pennant_write_gold(pennant_problem* pb, const char* base) {
    const int nz = pb->dm.numz, np = pb->dm.nump;
    std::vector<double> zr(nz), ze(nz), zp(nz), px(np), py(np);
    CK(pennant_get_field(pb, "zr", zr.data()));
    CK(pennant_get_field(pb, "ze", ze.data()));
    CK(pennant_get_field(pb, "zp", zp.data()));
    CK(hipMemcpy(px.data(), pb->f("px", 0), sizeof(double) * np, hipMemcpyDeviceToHost));
    CK(hipMemcpy(py.data(), pb->f("px", 1), sizeof(double) * np, hipMemcpyDeviceToHost));
    CK(copy_ctl_to_host(pb));
    return pennant_write_gold_host(base, pb->ctl_host->cycle, pb->ctl_host->time, zr.data(),
                                   ze.data(), zp.data(), pb->hm.znump.data(), nz, px.data(),
                                   py.data(), np, pb->hm.mapsp1.data());
}

}  // extern "C"
