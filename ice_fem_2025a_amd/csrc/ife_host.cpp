// ife_host.cpp -- host side of the engine: the C-ABI of include/ife.h, DOF numbering,
// CSR pattern, element->CSR scatter map, element colouring and device buffers.
//
// Reference correspondence (paths relative to the IceFEM source tree):
//   numbering   Physics/DofSpace.cpp:452-510 (per step, per field in DofNames order,
//               nodes ascending, dof = counter++) and getDofForNodes :974-1012
//   pattern     Physics/physics.cpp:228-251 + 70-78: the AIJ pattern is the union of the
//               element blocks the first assembly inserts (explicit zeros kept)
//   assembly    Physics/physics.cpp:255-303 (zero K/f, models in order, finalise)
//   history     Models/Fracture/PhaseField/PhaseField.cpp:123-163 (Commit / ResetStep)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <map>
#include <vector>

#include "ife_internal.h"

using namespace ife;

namespace {

thread_local std::string g_err;

struct ife_exception : std::runtime_error {
    int code;
    ife_exception(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define THROW(code, msg) throw ife_exception(code, msg)
#define HIPCHK(x)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) THROW(IFE_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class F>
int guard(F&& f) {
    try {
        f();
        return IFE_OK;
    } catch (ife_exception& e) {
        g_err = e.what();
        return e.code;
    } catch (std::bad_alloc&) {
        g_err = "host allocation failed";
        return IFE_ERR_NOMEM;
    } catch (std::exception& e) {
        g_err = e.what();
        return IFE_ERR_INVALID;
    }
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t n_) {
        free();
        n = n_;
        if (n_ == 0) return;
        hipError_t e = hipMalloc((void**)&p, n_ * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            THROW(IFE_ERR_NOMEM, std::string("hipMalloc of ") + std::to_string(n_ * sizeof(T)) + " bytes: " + hipGetErrorString(e));
        }
    }
    void upload(const T* h, size_t n_) {
        alloc(n_);
        if (n_) HIPCHK(hipMemcpy(p, h, n_ * sizeof(T), hipMemcpyHostToDevice));
    }
    void free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { free(); }
};

static void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)>& fn) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) {
        int64_t a = n * t / nt, b = n * (t + 1) / nt;
        th.emplace_back(fn, a, b);
    }
    for (auto& t : th) t.join();
}

enum { MT_PHASEFIELD = 0, MT_THERMAL = 1 };

struct Model {
    int type;
    std::string name;
    ife_material mat;
    ife_phasefield_opts pf;
    ife_model_opts opt;
    int Step_u = -1, Step_phase = -1, Step_T = 99;
    int field_u[6] = {-1, -1, -1, -1, -1, -1}, field_phase = -1, field_T = -1;
    int field_L = -1, field_L2 = -1;  // multiplier fields of the phase step (lMult_pf, lMult_pf2)
    int nfu = 0;        // u-step fields: dim, or 6 for the Cosserat model (u and micro-rotations)
    bool coss = false;  // Fracture/CossPhaseField
    bool HasT = false;
    int HistSize = 0;
    // device history [comp][E*nip]
    DevBuf<double> E, EOld, PS, PSOld, VD, VDOld, DVD, DVDOld, MH, MHOld;
    // per-element parameters (allocated at setup, used once set) and scalar overrides
    DevBuf<double> Gc_e, A_e;
    DevBuf<double> cossb;  // Cosserat per-IP tangent/stress coefficients [e][COSS_NF][nip]
    bool use_Gc_e = false, use_A_e = false;
    double n_override = std::numeric_limits<double>::quiet_NaN();
    double rm_eabs = -1, rm_erel = -1;
    int rm_maxit = -1, exact_tangent = 0;
};

struct Step {
    int64_t nrows = 0;
    int nf = 0;
    std::vector<int> fields;      // used fields of the step, DofNames order
    bool has_matrix = false;
    int64_t M = 0;                // nnz of one field's row block (every field pair present)
    // field pairs present: blk[fp][g] for row field fp and column field g.  Row block fp starts
    // at rb[fp], holds cnt[fp] column blocks per node row, column field g at block cpos[fp][g]
    std::vector<std::vector<char>> blk;
    std::vector<int64_t> rb;
    std::vector<int> cnt;
    std::vector<std::vector<int>> cpos;
    std::vector<int64_t> row_ptr;
    std::vector<int32_t> col;
    DevBuf<int64_t> d_row_ptr;
    DevBuf<int32_t> d_col;
    DevBuf<double> K, f, x[6];  // ife_state_kind order
    // Dirichlet constraints (Constrainer): constrained dofs ascending with their values, free
    // dofs ascending, the constrained pattern with each entry's offset in its source row
    bool has_cons = false;
    int64_t nfree = 0;
    std::vector<int64_t> c_row_ptr;
    std::vector<int32_t> c_col;
    DevBuf<int64_t> d_cdofs, d_free, d_crp, d_cfp;
    DevBuf<int32_t> d_ccol, d_rel, d_cfr;
    DevBuf<double> d_cvals, d_g, d_kcon, d_fcon, d_part, d_norms;
    KrylovState* kry = nullptr;  // ife_physics_solve's buffers (freed with the physics)
    DevBuf<LSDev> d_ls;          // the device line search's record
};

}  // namespace

namespace ife {
void set_last_error(const std::string& msg) { g_err = msg; }
}  // namespace ife

struct ife_physics {
    // mesh
    int dim = 0, ip1d = 3, n = 0, nip = 0;
    int64_t NN = 0, NE = 0;
    std::vector<double> coords, W;
    std::vector<int64_t> elems;
    std::vector<double> ext1d;  // [e][dim][3][3]
    // dofs
    std::vector<std::string> fname;
    std::vector<int> fstep;
    int nsteps = 0;
    std::vector<char> fused;     // field registered by a model
    std::vector<Model*> models;
    // plan
    bool planned = false, ready = false;
    std::vector<int32_t> rank;      // node -> group rank (-1 if unused)
    std::vector<int64_t> unique;    // group rank -> node
    int64_t Ng = 0;
    std::vector<int64_t> nbr_ptr;
    std::vector<int32_t> nbr;
    std::vector<int32_t> erank;     // [e][n]
    std::vector<uint8_t> loc;       // [e][n][n]
    std::vector<int64_t> color_ptr;
    std::vector<int32_t> color_elems;
    std::vector<int64_t> n2e_ptr;   // node (rank) -> adjacent elements
    std::vector<int32_t> n2e_elem;
    std::vector<uint8_t> n2e_loc;
    std::vector<Step> steps;
    std::vector<int> fpos;          // field -> position within its step (-1 unused)
    // device
    DevBuf<double> d_nodes, d_ext, d_lcrack, d_ipb, d_ebuf;
    // Cosserat u-step element records for crec_cap elements at a time (FP64-sized; mixed
    // mode's FP32 records fit twice as many); crec_chunk: elements per chunk set by the
    // "coss_record_chunk" parameter (0: as many as fit, negative: the coloured additions)
    DevBuf<double> d_crec;
    int64_t crec_cap = 0, crec_chunk = 0;
    // 2D element kernel in mixed mode ("elements_per_wave"): 1 one wavefront per element on the
    // VALU, 4 / 8 / 16 that many elements per wavefront on MFMA
    int epw2d = 1;
    DevBuf<int64_t> d_n2e_ptr;
    DevBuf<int32_t> d_n2e_elem;
    DevBuf<uint8_t> d_n2e_loc;
    DevBuf<int32_t> d_enode, d_erank, d_colors;
    DevBuf<uint8_t> d_loc;
    DevBuf<int64_t> d_nbr_ptr;
    DevBuf<Status> d_status;
    // time / precision
    double t = 0, dt = 1, du_dt = 1, ddu_dt = 1;
    int ts = IFE_TS_EULER;  // time scheme of UpdateVelAcc and its constants
    double ts_c0 = 1.0, ts_c1 = 0.0;
    int precision = IFE_PREC_FP64;
    // a phase step has been assembled: LCrack then holds its crack length, as the model's
    // LCrack member keeps the last Assemble_Phase's sum through other steps (PhaseField.cpp:540,650)
    bool phase_assembled = false;
    // launch timing
    bool profiling = false;
    struct Rec { hipEvent_t a, b; int64_t elems; };
    std::vector<Rec> recs, mat_recs, gat_recs;
    std::vector<hipEvent_t> event_pool;

    ~ife_physics() {
        for (auto* m : models) delete m;
        for (auto& r : recs) { event_pool.push_back(r.a); event_pool.push_back(r.b); }
        for (auto& r : mat_recs) { event_pool.push_back(r.a); event_pool.push_back(r.b); }
        for (auto& r : gat_recs) { event_pool.push_back(r.a); event_pool.push_back(r.b); }
        for (auto e : event_pool) (void)hipEventDestroy(e);
    }
    hipEvent_t get_event() {
        if (!event_pool.empty()) { hipEvent_t e = event_pool.back(); event_pool.pop_back(); return e; }
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) THROW(IFE_ERR_HIP, "hipEventCreate");
        return e;
    }
    int field_index(const char* nm) const {
        for (size_t i = 0; i < fname.size(); i++)
            if (fname[i] == nm) return (int)i;
        return -1;
    }
    Model* pf_model() const {
        for (auto* m : models)
            if (m->type == MT_PHASEFIELD) return m;
        return nullptr;
    }
    Model* th_model() const {
        for (auto* m : models)
            if (m->type == MT_THERMAL) return m;
        return nullptr;
    }
};

namespace {

// Factor a dense extraction operator into kron(Cz, kron(Cy, Cx)) (bezierExtraction3D.m:23)
// normalised with Cy[0][0] = Cz[0][0] = 1; false when C is not such a product.
bool factor_extraction(int dim, const double* C, double* out /* dim*9 */) {
    const int n = dim == 2 ? 9 : 27;
    double c00 = C[0];
    if (c00 == 0.0) return false;
    auto idx = [&](int x, int y, int z) { return x + 3 * y + 9 * z; };
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            out[0 * 9 + a * 3 + b] = C[idx(a, 0, 0) * n + idx(b, 0, 0)];
            out[1 * 9 + a * 3 + b] = C[idx(0, a, 0) * n + idx(0, b, 0)] / c00;
            if (dim == 3) out[2 * 9 + a * 3 + b] = C[idx(0, 0, a) * n + idx(0, 0, b)] / c00;
        }
    double mx = 0;
    for (int i = 0; i < n * n; i++) mx = std::max(mx, std::fabs(C[i]));
    for (int a = 0; a < n; a++)
        for (int b = 0; b < n; b++) {
            int ax = a % 3, ay = (a / 3) % 3, az = a / 9, bx = b % 3, by = (b / 3) % 3, bz = b / 9;
            double v = out[ax * 3 + bx] * out[9 + ay * 3 + by];
            if (dim == 3) v *= out[18 + az * 3 + bz];
            if (std::fabs(v - C[a * n + b]) > 1e-14 * mx) return false;
        }
    return true;
}

MatConst make_matconst(const ife_physics* ph, const Model* pfm, const Model* thm) {
    MatConst mc{};
    const Model* base = pfm ? pfm : thm;
    const ife_material& m = base->mat;
    mc.Young0 = m.Young; mc.Poisson = m.Poisson; mc.Density = m.Density; mc.Damping = m.Damping;
    mc.Rheology = m.Rheology;
    mc.HistSize = pfm ? pfm->HistSize : 0;
    mc.A0 = m.A; mc.nCreep = m.n; mc.PhiViscLim = m.PhiViscLim;
    mc.rm_eabs = -1; mc.rm_erel = -1; mc.rm_maxit = -1; mc.exact_tangent = 0;
    if (pfm) {
        if (!std::isnan(pfm->n_override)) mc.nCreep = pfm->n_override;
        mc.rm_eabs = pfm->rm_eabs; mc.rm_erel = pfm->rm_erel; mc.rm_maxit = pfm->rm_maxit;
        mc.exact_tangent = pfm->exact_tangent;
    }
    mc.l = m.l; mc.Gc_in = m.Gc; mc.pf_visc = m.Viscosity;
    mc.Split = m.Split;
    mc.TRef = m.TRef; mc.TOffset = m.TOffset; mc.QCreep = m.QCreep; mc.ft0 = m.ft0; mc.ftdeg = m.ftdeg;
    mc.sy0 = m.sy; mc.sydeg = m.sydeg;
    mc.B_dp = m.B_DP; mc.Ice_eRef = m.eRef; mc.Ice_cReduce = m.cReduce; mc.StarConvexGamma = m.StarConvexGamma;
    mc.l_coss = m.l_coss; mc.G_coss0 = m.G_coss; mc.hard_eRef = m.hard_eRef;  // CosseratMaterial.cpp:5-10
    if (m.Rheology == IFE_RHEO_VEP || m.Rheology == IFE_RHEO_DVEP) {  // SolidMaterial.cpp:49-64
        const double ya = M_PI / 180.0 * m.F_deg, pa = M_PI / 180.0 * m.G_deg;
        mc.yield_alpha = 6.0 * std::sin(ya) / (3.0 - std::sin(ya));
        mc.pot_alpha = 6.0 * std::sin(pa) / (3.0 - std::sin(pa));
        mc.hardening = m.hardening;
        mc.pl_visc = m.pl_visc;
    }
    // the thermal model reads its own material block (ThermalModel.cpp:61-62)
    if (thm) { mc.k = thm->mat.ThermalConductivity; mc.cp = thm->mat.HeatCapacity; mc.Density_th = thm->mat.Density; }
    if (pfm) {
        const ife_phasefield_opts& pf = pfm->pf;
        mc.ConsistentEnergy = pf.ConsistentEnergy;
        mc.DamFunction = pf.DamFunction; mc.DensityDamage = pf.DensityDamage; mc.GravityDamage = pf.GravityDamage;
        mc.ViscoDriving = pf.ViscoDriving; mc.ViscoDrivingFactor = pf.ViscoDrivingFactor; mc.k0 = pf.k0; mc.p = pf.p;
        mc.Distribution = pf.DistributionFunction; mc.ConsistentDegradation = pf.ConsistentDegradation;
        for (int i = 0; i < 5; i++) mc.coefs[i] = pf.Coefs[i];
        const ife_model_opts& o = pfm->opt;
        mc.ViscousStab = o.ViscousStab; mc.Irreversible = o.Irreversible; mc.Inertia = o.Inertia;
        mc.LumpedCap = o.LumpedPFCapacity; mc.LumpedDrive = o.LumpedDriving; mc.ViscousHeating = o.ViscousHeating;
        mc.porosity = o.porosity; mc.phase_T_from_dof0 = o.phase_T_from_dof0; mc.T_default = o.T_default;
        if (o.Gravity) {
            if (ph->dim == 3) { mc.g[0] = 0; mc.g[1] = 0; mc.g[2] = -9.81; }
            else { mc.g[0] = 0; mc.g[1] = -9.81; mc.g[2] = 0; }
        }
    }
    return mc;
}

void plan(ife_physics* ph) {
    if (ph->planned) return;
    if (ph->models.empty()) THROW(IFE_ERR_STATE, "no models added");
    const int n = ph->n;
    // ---- element group: unique nodes (elementgroup::GetUniqueNodes) -----------------
    ph->rank.assign(ph->NN, -1);
    for (int64_t i = 0; i < ph->NE * n; i++) {
        int64_t v = ph->elems[i];
        if (v < 0 || v >= ph->NN) THROW(IFE_ERR_INVALID, "element node id out of range");
        ph->rank[v] = 0;
    }
    ph->unique.clear();
    for (int64_t v = 0; v < ph->NN; v++)
        if (ph->rank[v] == 0) { ph->rank[v] = (int32_t)ph->unique.size(); ph->unique.push_back(v); }
    ph->Ng = (int64_t)ph->unique.size();
    const int64_t Ng = ph->Ng;
    ph->erank.resize(ph->NE * n);
    for (int64_t i = 0; i < ph->NE * n; i++) ph->erank[i] = ph->rank[ph->elems[i]];
    // ---- numbering per step (DofSpace.cpp:486-510) --------------------------------
    const int nfl = (int)ph->fname.size();
    ph->fpos.assign(nfl, -1);
    ph->steps.clear();
    ph->steps.resize(ph->nsteps);
    for (int s = 0; s < ph->nsteps; s++) {
        Step& st = ph->steps[s];
        for (int f = 0; f < nfl; f++)
            if (ph->fstep[f] == s && ph->fused[f]) { ph->fpos[f] = st.nf++; st.fields.push_back(f); }
        st.nrows = (int64_t)st.nf * Ng;
    }
    // ---- which steps receive element blocks, and the block must cover the step ----
    for (auto* m : ph->models) {
        auto need_exact = [&](int step, std::vector<int> flds) {
            Step& st = ph->steps[step];
            std::vector<int> a = st.fields, b = flds;
            std::sort(a.begin(), a.end());
            std::sort(b.begin(), b.end());
            if (a != b)
                THROW(IFE_ERR_UNSUPPORTED, "step " + std::to_string(step) +
                                               " mixes fields of different element blocks; the engine requires each "
                                               "matrix-assembling step to hold exactly one model block's fields");
            st.has_matrix = true;
        };
        if (m->type == MT_PHASEFIELD) {
            std::vector<int> fu;
            for (int d = 0; d < m->nfu; d++) fu.push_back(m->field_u[d]);
            need_exact(m->Step_u, fu);
            std::vector<int> fph = {m->field_phase};
            for (int fl : {m->field_L, m->field_L2})
                if (fl >= 0) fph.push_back(fl);
            need_exact(m->Step_phase, fph);
        } else {
            need_exact(m->Step_T, {m->field_T});
        }
    }
    // ---- node graph over group ranks: union of element node sets ------------------
    std::vector<int64_t> n2e_ptr(Ng + 1, 0);
    for (int64_t i = 0; i < ph->NE * n; i++) n2e_ptr[ph->erank[i] + 1]++;
    for (int64_t v = 0; v < Ng; v++) n2e_ptr[v + 1] += n2e_ptr[v];
    std::vector<int64_t> n2e(n2e_ptr[Ng]);
    ph->n2e_elem.resize(n2e_ptr[Ng]);
    ph->n2e_loc.resize(n2e_ptr[Ng]);
    {
        // elements in ascending index per node: the order the gather adds them in
        std::vector<int64_t> fill(n2e_ptr.begin(), n2e_ptr.end() - 1);
        for (int64_t e = 0; e < ph->NE; e++)
            for (int a = 0; a < n; a++) {
                int64_t k = fill[ph->erank[e * n + a]]++;
                n2e[k] = e;
                ph->n2e_elem[k] = (int32_t)e;
                ph->n2e_loc[k] = (uint8_t)a;
            }
    }
    ph->n2e_ptr = n2e_ptr;
    std::vector<std::vector<int32_t>> lists(Ng);
    parallel_for(Ng, [&](int64_t v0, int64_t v1) {
        std::vector<int64_t> mark(Ng, -1);
        for (int64_t v = v0; v < v1; v++) {
            auto& L = lists[v];
            for (int64_t k = n2e_ptr[v]; k < n2e_ptr[v + 1]; k++) {
                int64_t e = n2e[k];
                for (int a = 0; a < n; a++) {
                    int32_t w = ph->erank[e * n + a];
                    if (mark[w] != v) { mark[w] = v; L.push_back(w); }
                }
            }
            std::sort(L.begin(), L.end());
        }
    });
    ph->nbr_ptr.assign(Ng + 1, 0);
    for (int64_t v = 0; v < Ng; v++) ph->nbr_ptr[v + 1] = ph->nbr_ptr[v] + (int64_t)lists[v].size();
    ph->nbr.resize(ph->nbr_ptr[Ng]);
    for (int64_t v = 0; v < Ng; v++) std::copy(lists[v].begin(), lists[v].end(), ph->nbr.begin() + ph->nbr_ptr[v]);
    lists.clear();
    lists.shrink_to_fit();
    int64_t maxdeg = 0;
    for (int64_t v = 0; v < Ng; v++) maxdeg = std::max(maxdeg, ph->nbr_ptr[v + 1] - ph->nbr_ptr[v]);
    if (maxdeg > (ph->dim == 2 ? 64 : 128))
        THROW(IFE_ERR_UNSUPPORTED, "node degree above 128 in 3D or 64 in 2D (gather rows are staged in LDS)");
    // ---- per-step CSR: row (fp, v) = [g*Ng + nbr(v) for the fields g paired with fp] ----
    // Every field pair of a step is present except the two multipliers of AugmentedLagrangeMult2
    // and 3, whose blocks the phase step never inserts (PhaseField.cpp:717-728)
    const int64_t nnz_node = ph->nbr_ptr[Ng];
    for (int s = 0; s < ph->nsteps; s++) {
        Step& st = ph->steps[s];
        const int nf = st.nf;
        st.blk.assign(nf, std::vector<char>(nf, 1));
        for (auto* m : ph->models)
            if (m->type == MT_PHASEFIELD && m->Step_phase == s && m->field_L >= 0 && m->field_L2 >= 0) {
                st.blk[ph->fpos[m->field_L]][ph->fpos[m->field_L2]] = 0;
                st.blk[ph->fpos[m->field_L2]][ph->fpos[m->field_L]] = 0;
            }
        st.rb.assign(nf, 0);
        st.cnt.assign(nf, 0);
        st.cpos.assign(nf, std::vector<int>(nf, -1));
        int64_t off = 0;
        for (int fp = 0; fp < nf; fp++) {
            for (int g = 0; g < nf; g++)
                if (st.blk[fp][g]) st.cpos[fp][g] = st.cnt[fp]++;
            st.rb[fp] = off;
            off += (int64_t)st.cnt[fp] * nnz_node;
        }
        st.row_ptr.assign(st.nrows + 1, 0);
        if (!st.has_matrix) { st.M = 0; continue; }
        st.M = (int64_t)nf * nnz_node;
        for (int fp = 0; fp < nf; fp++)
            for (int64_t v = 0; v <= Ng; v++) st.row_ptr[(int64_t)fp * Ng + v] = st.rb[fp] + (int64_t)st.cnt[fp] * ph->nbr_ptr[v];
        st.row_ptr[st.nrows] = off;
        st.col.resize((size_t)off);
        parallel_for(Ng, [&](int64_t v0, int64_t v1) {
            for (int fp = 0; fp < nf; fp++)
                for (int64_t v = v0; v < v1; v++) {
                    int64_t deg = ph->nbr_ptr[v + 1] - ph->nbr_ptr[v];
                    int32_t* out = st.col.data() + st.row_ptr[(int64_t)fp * Ng + v];
                    for (int g = 0; g < nf; g++) {
                        if (!st.blk[fp][g]) continue;
                        const int c = st.cpos[fp][g];
                        for (int64_t k = 0; k < deg; k++) out[c * deg + k] = (int32_t)(g * Ng + ph->nbr[ph->nbr_ptr[v] + k]);
                    }
                }
        });
    }
    // ---- scatter slots: loc[e][a][b] = position of node b in node a's neighbour list
    ph->loc.resize((size_t)ph->NE * n * n);
    parallel_for(ph->NE, [&](int64_t e0, int64_t e1) {
        for (int64_t e = e0; e < e1; e++)
            for (int a = 0; a < n; a++) {
                int32_t ra = ph->erank[e * n + a];
                const int32_t* b0 = ph->nbr.data() + ph->nbr_ptr[ra];
                const int32_t* b1 = ph->nbr.data() + ph->nbr_ptr[ra + 1];
                for (int b = 0; b < n; b++) {
                    const int32_t* it = std::lower_bound(b0, b1, ph->erank[e * n + b]);
                    ph->loc[((size_t)e * n + a) * n + b] = (uint8_t)(it - b0);
                }
            }
    });
    // ---- greedy first-fit colouring: no two elements of a colour share a node -----
    std::vector<uint64_t> used(Ng, 0);
    std::vector<int> color(ph->NE);
    int ncol = 0;
    for (int64_t e = 0; e < ph->NE; e++) {
        uint64_t m = 0;
        for (int a = 0; a < n; a++) m |= used[ph->erank[e * n + a]];
        if (m == ~0ull) THROW(IFE_ERR_UNSUPPORTED, "element colouring needs more than 64 colours");
        int c = __builtin_ctzll(~m);
        color[e] = c;
        ncol = std::max(ncol, c + 1);
        for (int a = 0; a < n; a++) used[ph->erank[e * n + a]] |= (1ull << c);
    }
    ph->color_ptr.assign(ncol + 1, 0);
    for (int64_t e = 0; e < ph->NE; e++) ph->color_ptr[color[e] + 1]++;
    for (int c = 0; c < ncol; c++) ph->color_ptr[c + 1] += ph->color_ptr[c];
    ph->color_elems.resize(ph->NE);
    {
        std::vector<int64_t> fill(ph->color_ptr.begin(), ph->color_ptr.end() - 1);
        for (int64_t e = 0; e < ph->NE; e++) ph->color_elems[fill[color[e]]++] = (int32_t)e;
    }
    ph->planned = true;
}

void setup_device(ife_physics* ph) {
    plan(ph);
    if (ph->ready) return;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev < 1) THROW(IFE_ERR_HIP, "no HIP device");
    const int n = ph->n;
    std::vector<double> nodes4((size_t)ph->NN * 4, 0.0);
    for (int64_t v = 0; v < ph->NN; v++) {
        for (int d = 0; d < ph->dim; d++) nodes4[v * 4 + d] = ph->coords[v * ph->dim + d];
        nodes4[v * 4 + 3] = ph->W[v];
    }
    ph->d_nodes.upload(nodes4.data(), nodes4.size());
    std::vector<int32_t> en(ph->NE * n);
    for (int64_t i = 0; i < ph->NE * n; i++) en[i] = (int32_t)ph->elems[i];
    ph->d_enode.upload(en.data(), en.size());
    ph->d_erank.upload(ph->erank.data(), ph->erank.size());
    ph->d_loc.upload(ph->loc.data(), ph->loc.size());
    ph->d_ext.upload(ph->ext1d.data(), ph->ext1d.size());
    ph->d_nbr_ptr.upload(ph->nbr_ptr.data(), ph->nbr_ptr.size());
    ph->d_colors.upload(ph->color_elems.data(), ph->color_elems.size());
    ph->d_status.alloc(1);
    HIPCHK(hipMemset(ph->d_status.p, 0, sizeof(Status)));
    // 2D stores the 29 rows its element kernel reads (W2C_NF in ife_kernels.hip), 3D all 70
    ph->d_ipb.alloc((size_t)coefficient_rows(ph->dim) * ph->NE * ph->nip);
    ph->d_ebuf.alloc((size_t)ph->NE * ph->n * record_size(ph->dim));
    ph->d_n2e_ptr.upload(ph->n2e_ptr.data(), ph->n2e_ptr.size());
    ph->d_n2e_elem.upload(ph->n2e_elem.data(), ph->n2e_elem.size());
    ph->d_n2e_loc.upload(ph->n2e_loc.data(), ph->n2e_loc.size());
    ph->d_lcrack.alloc(ph->NE);
    HIPCHK(hipMemset(ph->d_lcrack.p, 0, ph->NE * sizeof(double)));
    for (auto& st : ph->steps) {
        st.d_row_ptr.upload(st.row_ptr.data(), st.row_ptr.size());
        st.d_col.upload(st.col.data(), st.col.size());
        st.K.alloc(std::max<size_t>(st.col.size(), 1));
        st.f.alloc(std::max<int64_t>(st.nrows, 1));
        HIPCHK(hipMemset(st.K.p, 0, st.K.n * sizeof(double)));
        HIPCHK(hipMemset(st.f.p, 0, st.f.n * sizeof(double)));
        for (int k = 0; k < 6; k++) {
            st.x[k].alloc(std::max<int64_t>(st.nrows, 1));
            HIPCHK(hipMemset(st.x[k].p, 0, st.x[k].n * sizeof(double)));
        }
    }
    const size_t NIPT = (size_t)ph->NE * ph->nip;
    for (auto* m : ph->models) {
        if (m->type != MT_PHASEFIELD) continue;
        size_t hs = (size_t)std::max(m->HistSize, 1);
        DevBuf<double>* one[] = {&m->E, &m->EOld, &m->VD, &m->VDOld, &m->DVD, &m->DVDOld};
        for (auto* b : one) b->alloc(NIPT);
        const size_t nps = m->coss ? 18 : 6;
        m->PS.alloc(nps * NIPT);
        m->PSOld.alloc(nps * NIPT);
        if (m->coss) m->cossb.alloc((size_t)COSS_NF * NIPT);
        m->MH.alloc(hs * NIPT);
        m->MHOld.alloc(hs * NIPT);
        m->Gc_e.alloc(ph->NE);
        m->A_e.alloc(ph->NE);
        DevBuf<double>* all[] = {&m->E, &m->EOld, &m->VD, &m->VDOld, &m->DVD, &m->DVDOld, &m->PS, &m->PSOld, &m->MH, &m->MHOld};
        for (auto* b : all) HIPCHK(hipMemset(b->p, 0, b->n * sizeof(double)));
    }
    // Cosserat u-step records: up to 64 GiB, leaving 16 GiB of the device free; without room
    // for 256 elements the u step falls back to the coloured additions into the CSR
    bool any_coss = false;
    for (auto* m : ph->models) any_coss = any_coss || (m->type == MT_PHASEFIELD && m->coss);
    if (any_coss) {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        const size_t per = (size_t)ph->n * coss_record_size(ph->dim) * sizeof(double);
        const size_t reserve = (size_t)16 << 30, cap = (size_t)64 << 30;
        const size_t budget = fr > reserve ? std::min(fr - reserve, cap) : 0;
        const int64_t ne = std::min<int64_t>(ph->NE, (int64_t)(budget / per));
        if (ne >= std::min<int64_t>(ph->NE, 256)) {
            ph->d_crec.alloc((size_t)ne * ph->n * coss_record_size(ph->dim));
            ph->crec_cap = ne;
        }
    }
    HIPCHK(hipDeviceSynchronize());
    ph->ready = true;
}

bool is_ho_distribution(int method) {  // NeedsGrad2 (PhaseFieldUtil.cpp:18-22)
    return method == IFE_DIST_HO_LINEAR || method == IFE_DIST_HO_LINEAR2 || method == IFE_DIST_HO_CZM ||
           method == IFE_DIST_HO_QUADRATIC || method == IFE_DIST_CUSTOM;
}

void require_ready(const ife_physics* ph) {
    if (!ph) THROW(IFE_ERR_INVALID, "null physics");
    if (!ph->ready) THROW(IFE_ERR_STATE, "ife_physics_setup has not been called");
}

// Bernstein basis of degree 2 at x with its first and second derivatives
// (BaseElemType.cpp:197-216) and the 3-point Gauss
// rule on [0,1] (IPScheme.cpp:62-109)
void bernstein_tables(KParams& p) {
    const double xs[3] = {(-std::sqrt(3.0 / 5.0) + 1) / 2, (0.0 + 1) / 2, (std::sqrt(3.0 / 5.0) + 1) / 2};
    const double ws[3] = {5.0 / 9.0 / 2, 8.0 / 9.0 / 2, 5.0 / 9.0 / 2};
    const int pdeg = 2;
    const int fact[3] = {1, 1, 2};
    for (int q = 0; q < 3; q++) {
        double x = xs[q];
        for (int i = 0; i < 3; i++) {
            double pre = (double)(fact[pdeg] / fact[pdeg - i] / fact[i]);
            p.B1[q][i] = pre * std::pow(1.0 - x, std::max(0, pdeg - i)) * std::pow(x, std::max(0, i));
            p.dB1[q][i] = pre * ((i - pdeg) * std::pow(1.0 - x, std::max(0, pdeg - i - 1)) * std::pow(x, std::max(0, i)) +
                                 i * std::pow(1.0 - x, pdeg - i) * std::pow(x, std::max(0, i - 1)));
            // second derivatives of (1-x)^2, 2x(1-x), x^2
            p.d2B1[q][i] = i == 1 ? -4.0 : 2.0;
        }
        p.wq[q] = ws[q];
    }
}

KParams base_params(ife_physics* ph) {
    KParams p{};
    p.nodes = ph->d_nodes.p;
    p.enode = ph->d_enode.p;
    p.erank = ph->d_erank.p;
    p.loc = ph->d_loc.p;
    p.ext = ph->d_ext.p;
    p.nbr_ptr = ph->d_nbr_ptr.p;
    p.Ng = ph->Ng;
    p.NE = ph->NE;
    bernstein_tables(p);
    Model* pfm = ph->pf_model();
    Model* thm = ph->th_model();
    p.mc = make_matconst(ph, pfm, thm);
    p.t = ph->t; p.dt = ph->dt; p.du_dt = ph->du_dt; p.ddu_dt = ph->ddu_dt;
    p.has_pf = pfm != nullptr;
    p.has_th = thm != nullptr;
    p.has_T = (pfm && pfm->HasT) || thm != nullptr;
    auto stepio = [&](int s) {
        StepIO io{};
        if (s < 0 || s >= ph->nsteps) return io;
        Step& st = ph->steps[s];
        io.x = st.x[IFE_STATE].p; io.xold = st.x[IFE_STATE_OLD].p; io.dx = st.x[IFE_DSTATE].p; io.ddx = st.x[IFE_DDSTATE].p;
        io.K = st.K.p; io.f = st.f.p; io.M = st.M; io.nf = st.nf;
        return io;
    };
    if (pfm) {
        p.su = stepio(pfm->Step_u);
        p.sph = stepio(pfm->Step_phase);
        for (int d = 0; d < pfm->nfu; d++) p.fpos_u[d] = ph->fpos[pfm->field_u[d]];
        p.coss = pfm->coss;
        p.cossb = pfm->cossb.p;
        p.fpos_ph = ph->fpos[pfm->field_phase];
        // the phase step's rows of phi and of the multipliers (absent fields: -1)
        const Step& sp = ph->steps[pfm->Step_phase];
        const int lf[3] = {pfm->field_phase, pfm->field_L, pfm->field_L2};
        p.lag = pfm->field_L >= 0 ? pfm->opt.Irreversible : 0;
        for (int i = 0; i < 3; i++) {
            const int fi = lf[i] >= 0 ? ph->fpos[lf[i]] : -1;
            p.ph_rb[i] = fi >= 0 ? sp.rb[fi] : 0;
            p.ph_cnt[i] = fi >= 0 ? sp.cnt[fi] : 0;
            p.ph_fpos[i] = fi;
            for (int j = 0; j < 3; j++) {
                const int fj = lf[j] >= 0 ? ph->fpos[lf[j]] : -1;
                p.ph_col[i][j] = fi >= 0 && fj >= 0 ? sp.cpos[fi][fj] : -1;
            }
        }
        p.E = pfm->E.p; p.EOld = pfm->EOld.p; p.PS = pfm->PS.p; p.PSOld = pfm->PSOld.p;
        p.VD = pfm->VD.p; p.VDOld = pfm->VDOld.p; p.DVD = pfm->DVD.p; p.DVDOld = pfm->DVDOld.p;
        p.MH = pfm->MH.p; p.MHOld = pfm->MHOld.p;
        p.Gc_e = pfm->use_Gc_e ? pfm->Gc_e.p : nullptr;
        p.A_e = pfm->use_A_e ? pfm->A_e.p : nullptr;
    }
    int fT = -1;
    if (thm) fT = thm->field_T;
    else if (pfm && pfm->HasT) fT = pfm->field_T;
    if (fT >= 0) {
        p.sT = stepio(ph->fstep[fT]);
        p.fpos_T = ph->fpos[fT];
        p.fpos_Tth = ph->fpos[fT];
    }
    p.NIPT = ph->NE * ph->nip;
    p.st = ph->d_status.p;
    p.lcrack_elem = ph->d_lcrack.p;
    p.ipb = ph->d_ipb.p;
    p.ebuf = ph->d_ebuf.p;
    p.n2e_ptr = ph->d_n2e_ptr.p;
    p.n2e_elem = ph->d_n2e_elem.p;
    p.n2e_loc = ph->d_n2e_loc.p;
    p.mixed = ph->precision == IFE_PREC_MIXED;
    p.epw2d = ph->epw2d;
    return p;
}

// KParams of one assembly of every step into the engine-owned K/f (what run(all) launches)
KParams all_steps_params(ife_physics* ph) {
    require_ready(ph);
    KParams p = base_params(ph);
    Model* pfm = ph->pf_model();
    Model* thm = ph->th_model();
    p.do_u = pfm != nullptr;
    p.do_ph = pfm != nullptr;
    p.do_Tpf = pfm && pfm->HasT && pfm->opt.ViscousHeating;
    p.do_th = thm != nullptr;
    p.elist = nullptr;
    p.nlist = ph->NE;
    return p;
}

void run(ife_physics* ph, bool all, int step, double* const* Kv, double* const* fv, hipStream_t s,
         const int* gate = nullptr) {
    require_ready(ph);
    KParams p = base_params(ph);
    p.gate = gate;
    Model* pfm = ph->pf_model();
    Model* thm = ph->th_model();
    std::vector<int> which;
    if (all) for (int i = 0; i < ph->nsteps; i++) which.push_back(i);
    else {
        if (step < 0 || step >= ph->nsteps) THROW(IFE_ERR_INVALID, "step out of range");
        which.push_back(step);
    }
    auto in = [&](int st) { return std::find(which.begin(), which.end(), st) != which.end(); };
    // caller buffers override engine-owned ones
    for (size_t i = 0; i < which.size(); i++) {
        int st = which[i];
        double* K = Kv ? Kv[i] : nullptr;
        double* f = fv ? fv[i] : nullptr;
        if (!K) K = ph->steps[st].K.p;
        if (!f) f = ph->steps[st].f.p;
        if (pfm && pfm->Step_u == st) { p.su.K = K; p.su.f = f; }
        if (pfm && pfm->Step_phase == st) { p.sph.K = K; p.sph.f = f; }
        if (p.has_T && ph->fstep[thm ? thm->field_T : pfm->field_T] == st) { p.sT.K = K; p.sT.f = f; }
    }
    // Physics::Assemble zeroes f and K and the models add into them (physics.cpp:260-284);
    // here the gather kernel writes every entry of every node row exactly once instead.
    p.do_u = pfm && in(pfm->Step_u);
    p.do_ph = pfm && in(pfm->Step_phase);
    p.do_Tpf = pfm && pfm->HasT && pfm->opt.ViscousHeating && in(pfm->Step_T);
    p.do_th = thm && in(thm->Step_T);
    if (p.do_ph) ph->phase_assembled = true;
    HIPCHK(launch_zero_status(ph->d_status.p, s, gate));
    if (!(p.do_u || p.do_ph || p.do_Tpf || p.do_th)) return;
    // material update over all IPs, element records for all elements in one launch, then
    // the per-node gather into the CSR rows
    p.elist = nullptr;
    p.nlist = ph->NE;
    auto timed = [&](std::vector<ife_physics::Rec>& into, int64_t elems, const std::function<hipError_t()>& launch) {
        if (ph->profiling) {
            ife_physics::Rec r{ph->get_event(), ph->get_event(), elems};
            HIPCHK(hipEventRecord(r.a, s));
            HIPCHK(launch());
            HIPCHK(hipEventRecord(r.b, s));
            into.push_back(r);
        } else {
            HIPCHK(launch());
        }
    };
    // higher-order crack densities: their derivative parts go into the phase records
    // between the element kernel and the gather
    const bool ho = p.do_ph && is_ho_distribution(pfm->pf.DistributionFunction);
    if (pfm && pfm->coss) {
        // Cosserat: the material kernel updates every IP; the u step's 162-dof element
        // blocks go straight into the zeroed CSR values colour by colour; the phase and T
        // steps use the shared element and gather kernels
        timed(ph->mat_recs, 0, [&] { return launch_coss_material(p, ph->dim, s); });
        if (p.do_u && ph->crec_cap > 0 && ph->crec_chunk >= 0) {
            // element records of a chunk of elements, then the rows of the chunk's nodes,
            // chunk after chunk in element order (FP32 records: twice the elements per chunk)
            int64_t chunk = ph->crec_cap * (p.mixed ? 2 : 1);
            if (ph->crec_chunk > 0) chunk = std::min(chunk, ph->crec_chunk);
            for (int64_t e0 = 0; e0 < ph->NE; e0 += chunk) {
                const int64_t e1 = std::min(ph->NE, e0 + chunk);
                KParams q = p;
                q.elist = nullptr;
                q.nlist = e1 - e0;
                q.crec = ph->d_crec.p;
                q.crec_e0 = e0;
                timed(ph->recs, q.nlist, [&] { return launch_coss_assemble(q, ph->dim, s); });
                timed(ph->gat_recs, 0, [&] { return launch_coss_gather(q, ph->dim, e0, e1, s); });
            }
        } else if (p.do_u) {
            const Step& su = ph->steps[pfm->Step_u];
            HIPCHK(launch_fill(p.su.K, (int64_t)su.col.size(), 0.0, s, gate));
            HIPCHK(launch_fill(p.su.f, su.nrows, 0.0, s, gate));
            for (size_t c = 0; c + 1 < ph->color_ptr.size(); c++) {
                KParams q = p;
                q.elist = ph->d_colors.p + ph->color_ptr[c];
                q.nlist = ph->color_ptr[c + 1] - ph->color_ptr[c];
                timed(ph->recs, q.nlist, [&] { return launch_coss_assemble(q, ph->dim, s); });
            }
        }
        KParams q = p;
        q.do_u = 0;
        if (q.do_ph || q.do_Tpf || q.do_th) {
            HIPCHK(launch_assemble(q, ph->dim, s));
            if (ho) HIPCHK(launch_ho_phase(q, ph->dim, s));
            timed(ph->gat_recs, 0, [&] { return launch_gather(q, ph->dim, s); });
        }
        return;
    }
    timed(ph->mat_recs, 0, [&] { return launch_material(p, ph->dim, s); });
    timed(ph->recs, ph->NE, [&] { return launch_assemble(p, ph->dim, s); });
    if (ho) HIPCHK(launch_ho_phase(p, ph->dim, s));
    timed(ph->gat_recs, 0, [&] { return launch_gather(p, ph->dim, s); });
}

// Adjoint of the full assembly's residuals (u, phase and T steps of ife_physics_assemble_all)
// with respect to the per-element Gc, the per-element creep prefactor A and the Glen exponent
// n, contracted with the cotangents lam[step].  Recomputes every IP's state from the current
// state vectors and the committed history; writes no history.
void run_adjoint(ife_physics* ph, const double* const* lam, double* g_Gc, double* g_A, double* g_n, hipStream_t s) {
    if (!ph) THROW(IFE_ERR_INVALID, "null physics");
    Model* pfm = ph->pf_model();
    if (!pfm) THROW(IFE_ERR_STATE, "the adjoint needs a Fracture/PhaseField model");
    const ife_material& mat = pfm->mat;
    const int rh = mat.Rheology;
    if (rh != IFE_RHEO_LE && rh != IFE_RHEO_VE && rh != IFE_RHEO_DVE)
        THROW(IFE_ERR_UNSUPPORTED, "adjoint: rheology must be LinearElastic, ViscoElastic or DamViscoElastic");
    if (mat.Split != IFE_SPLIT_NONE && mat.Split != IFE_SPLIT_VOLSTRAINS && mat.Split != IFE_SPLIT_VOLSTRESS)
        THROW(IFE_ERR_UNSUPPORTED, "adjoint: energy split must be None, VolStrains or VolStress");
    if (pfm->opt.LumpedDriving || pfm->opt.LumpedPFCapacity)
        THROW(IFE_ERR_UNSUPPORTED, "adjoint: lumped driving / lumped capacity are not differentiated");
    if (pfm->pf.ViscoDrivingFactor > 1.0e-9) THROW(IFE_ERR_UNSUPPORTED, "adjoint: viscous driving is not differentiated");
    if (is_ho_distribution(pfm->pf.DistributionFunction))
        THROW(IFE_ERR_UNSUPPORTED, "adjoint: the higher-order crack densities are not differentiated");
    require_ready(ph);
    KParams p = base_params(ph);
    p.do_u = 1;
    p.do_ph = 1;
    p.do_Tpf = pfm->HasT && pfm->opt.ViscousHeating;
    p.do_th = 0;
    auto pick = [&](int st) -> const double* { return (lam && st >= 0 && st < ph->nsteps) ? lam[st] : nullptr; };
    p.lam_u = pick(pfm->Step_u);
    p.lam_ph = pick(pfm->Step_phase);
    p.lam_T = p.do_Tpf ? pick(pfm->Step_T) : nullptr;
    p.g_Gc = g_Gc;
    p.g_A = g_A;
    p.g_n = g_n;
    HIPCHK(launch_zero_status(ph->d_status.p, s));
    HIPCHK(launch_adjoint(p, ph->dim, s));
}

DevBuf<double>* history_buf(Model* m, const std::string& nm, int64_t& ncomp) {
    ncomp = 1;
    if (nm == "Energy") return &m->E;
    if (nm == "EnergyOld") return &m->EOld;
    if (nm == "ViscDiss") return &m->VD;
    if (nm == "ViscDissOld") return &m->VDOld;
    if (nm == "DamViscDiss") return &m->DVD;
    if (nm == "DamViscDissOld") return &m->DVDOld;
    if (nm == "PlasticStrain") { ncomp = m->coss ? 18 : 6; return &m->PS; }
    if (nm == "PlasticStrainOld") { ncomp = m->coss ? 18 : 6; return &m->PSOld; }
    if (nm == "MatHist") { ncomp = std::max(m->HistSize, 1); return &m->MH; }
    if (nm == "MatHistOld") { ncomp = std::max(m->HistSize, 1); return &m->MHOld; }
    return nullptr;
}

}  // namespace

// ==================================================================================
// Batches of independent domains
// ==================================================================================
struct ife_batch {
    std::vector<ife_physics*> doms;
    int dim = 0, mixed = 0, vep = 0, exact = 0, full = 0, lin = 0;
    DevBuf<KParams> dP;
    DevBuf<int64_t> bp[3];      // block prefix sums: material, element, gather kernels
    int64_t tot[3] = {0, 0, 0};
    int64_t nodes = 0;          // element-group nodes over all domains
    int64_t grid[3] = {0, 0, 0};  // workgroups per launch (0: one per block)
    int nta = 256;  // lanes per element of the element kernel (64 in 2D: one wavefront)
    // 2D element kernel: elements per wavefront (1: assemble2d_wave; 4, 8, 16: assemble2d_mfma,
    // mixed mode), and the block tables of the MFMA forms (ceil(elements / epw) per domain)
    int epw = 1;
    DevBuf<int64_t> bpw[3];
    int64_t totw[3] = {0, 0, 0};
    hipStream_t cap = nullptr;    // capture stream (default priority)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~ife_batch() {
        drop_graph();
        if (cap) (void)hipStreamDestroy(cap);
    }
    BatchArgs args() const {
        BatchArgs a{};
        a.P = dP.p;
        a.nd = (int)doms.size();
        a.dim = dim; a.mixed = mixed; a.vep = vep; a.exact = exact; a.full = full; a.lin = lin;
        for (int k = 0; k < 3; k++) { a.bptr[k] = bp[k].p; a.tot[k] = tot[k]; }
        set_epw(a, epw);
        return a;
    }
    void set_epw(BatchArgs& a, int e) const {
        a.epw = e;
        const int k = e == 4 ? 0 : e == 8 ? 1 : e == 16 ? 2 : -1;
        a.bptr_w = k >= 0 ? bpw[k].p : nullptr;
        a.tot_w = k >= 0 ? totw[k] : 0;
    }
};

namespace {

// (Re)read every domain's parameters (time, material constants, precision) into the
// device parameter array; the block tables depend only on the meshes.
void batch_update(ife_batch* b) {
    std::vector<KParams> P;
    int vep = -1, exact = -1, full = 0, mixed = -1, lin = 1;
    for (auto* ph : b->doms) {
        KParams p = all_steps_params(ph);
        lin &= p.mc.Rheology == IFE_RHEO_LE;
        const int v = p.mc.Rheology == IFE_RHEO_VEP;
        const int x = !v && p.mc.exact_tangent;
        const int sp = p.mc.Split;
        full |= !(sp == IFE_SPLIT_NONE || sp == IFE_SPLIT_VOLSTRAINS || sp == IFE_SPLIT_VOLSTRESS);
        if (vep < 0) { vep = v; exact = x; mixed = p.mixed; }
        if (v != vep || x != exact)
            THROW(IFE_ERR_UNSUPPORTED, "batch: domains need the same return map (Glen / visco-elasto-plastic, tangent)");
        if (p.mixed != mixed) THROW(IFE_ERR_UNSUPPORTED, "batch: domains need the same precision");
        if (p.has_pf && is_ho_distribution(p.mc.Distribution))
            THROW(IFE_ERR_UNSUPPORTED, "batch: the higher-order crack densities run through ife_physics_assemble");
        P.push_back(p);
    }
    b->vep = vep; b->exact = exact; b->full = full; b->mixed = mixed; b->lin = lin;
    b->dP.upload(P.data(), P.size());
    b->drop_graph();  // kernel arguments point at dP, which upload may have moved
}

void batch_enqueue(ife_batch* b, hipStream_t s) {
    const BatchArgs a = b->args();
    HIPCHK(launch_batch_zero(a, s));
    HIPCHK(launch_batch_material(a, b->grid[0], s));
    HIPCHK(launch_batch_assemble(a, b->grid[1], b->nta, s));
    HIPCHK(launch_batch_gather(a, b->grid[2], s));
    for (auto* ph : b->doms) ph->phase_assembled |= ph->pf_model() != nullptr;
}

void batch_capture(ife_batch* b) {
    b->drop_graph();
    if (!b->cap) HIPCHK(hipStreamCreate(&b->cap));
    HIPCHK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    try {
        batch_enqueue(b, b->cap);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(b->cap, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    HIPCHK(hipStreamEndCapture(b->cap, &b->graph));
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
}

// Launch-shape autotuner: per kernel, the median of three timed launches of every
// candidate shape (after one warm-up launch) on the current inputs; the fastest is kept.
void batch_tune(ife_batch* b, hipStream_t s, ife_batch_tuning* out) {
    batch_enqueue(b, s);  // every kernel's inputs in place
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int timed = 0;
    auto time_it = [&](const std::function<hipError_t()>& launch) {
        HIPCHK(launch());
        float t[3];
        for (int r = 0; r < 3; r++) {
            HIPCHK(hipEventRecord(e0, s));
            HIPCHK(launch());
            HIPCHK(hipEventRecord(e1, s));
            HIPCHK(hipEventSynchronize(e1));
            HIPCHK(hipEventElapsedTime(&t[r], e0, e1));
        }
        std::sort(t, t + 3);
        timed++;
        return (double)t[1];
    };
    auto candidates = [&](int64_t total) {
        std::vector<int64_t> c = {0};
        for (int64_t g = 256; g < total && g <= 256 * 64; g *= 2) c.push_back(g);
        return c;
    };
    const BatchArgs a = b->args();
    double best[3] = {1e300, 1e300, 1e300};
    try {
        for (int64_t g : candidates(b->tot[0])) {
            double ms = time_it([&] { return launch_batch_material(a, g, s); });
            if (ms < best[0]) { best[0] = ms; b->grid[0] = g; }
        }
        // element kernel: 3D one 256-lane workgroup per element, 2D one wavefront per element
        // (four per workgroup); the candidates are grid shapes in workgroups.  2D mixed mode
        // also times the MFMA kernel with 4, 8 and 16 elements per wavefront (one launch
        // wavefront per group of elements).
        const int nt = b->dim == 2 ? 64 : 256;
        const int64_t wgs = b->dim == 2 ? (b->tot[1] + 3) / 4 : b->tot[1];
        BatchArgs aw = a;
        b->set_epw(aw, 1);
        for (int64_t g : candidates(wgs)) {
            double ms = time_it([&] { return launch_batch_assemble(aw, g, nt, s); });
            if (ms < best[1]) { best[1] = ms; b->grid[1] = g; b->nta = nt; b->epw = 1; }
        }
        if (b->dim == 2 && b->mixed) {
            for (int e : {4, 8, 16}) {
                BatchArgs am = a;
                b->set_epw(am, e);
                double ms = time_it([&] { return launch_batch_assemble(am, 0, nt, s); });
                if (ms < best[1]) { best[1] = ms; b->grid[1] = 0; b->nta = nt; b->epw = e; }
            }
        }
        for (int64_t g : candidates(b->tot[2])) {
            double ms = time_it([&] { return launch_batch_gather(a, g, s); });
            if (ms < best[2]) { best[2] = ms; b->grid[2] = g; }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->drop_graph();
    if (out) {
        out->grid_material = b->grid[0];
        out->grid_assemble = b->grid[1];
        out->grid_gather = b->grid[2];
        out->lanes_assemble = b->nta;
        out->elems_per_wave = b->epw;
        out->candidates_timed = timed;
        out->ms_material = best[0];
        out->ms_assemble = best[1];
        out->ms_gather = best[2];
    }
}

}  // namespace

// ==================================================================================
// C-ABI
// ==================================================================================
extern "C" {

const char* ife_last_error(void) { return g_err.c_str(); }
int ife_api_version(void) { return IFE_API_VERSION; }

int ife_physics_create(const ife_mesh_desc* mesh, const ife_dofs_desc* dofs, ife_physics** out) {
    return guard([&] {
        if (!mesh || !dofs || !out) THROW(IFE_ERR_INVALID, "null argument");
        *out = nullptr;
        if (!((mesh->dim == 2 && mesh->nodes_per_elem == 9) || (mesh->dim == 3 && mesh->nodes_per_elem == 27)))
            THROW(IFE_ERR_UNSUPPORTED, "element type: dim 2 needs 9 nodes (B2_2), dim 3 needs 27 (NURBS3_Cube2)");
        if (mesh->ipcount1D != 3) THROW(IFE_ERR_UNSUPPORTED, "ipcount1D must be 3");
        if (mesh->n_nodes <= 0 || mesh->n_elems <= 0 || !mesh->coords || !mesh->elems)
            THROW(IFE_ERR_INVALID, "empty mesh");
        if (mesh->n_elems > INT32_MAX || mesh->n_nodes > INT32_MAX) THROW(IFE_ERR_UNSUPPORTED, "mesh too large for 32-bit ids");
        if (!mesh->extraction && !mesh->extraction_1d) THROW(IFE_ERR_INVALID, "no extraction operators");
        std::unique_ptr<ife_physics> ph(new ife_physics());
        ph->dim = mesh->dim;
        ph->n = mesh->nodes_per_elem;
        ph->nip = ph->n;
        ph->NN = mesh->n_nodes;
        ph->NE = mesh->n_elems;
        ph->coords.assign(mesh->coords, mesh->coords + ph->NN * ph->dim);
        ph->W.assign(ph->NN, 1.0);
        if (mesh->weights && ph->dim == 3) ph->W.assign(mesh->weights, mesh->weights + ph->NN);
        ph->elems.assign(mesh->elems, mesh->elems + ph->NE * ph->n);
        ph->ext1d.resize((size_t)ph->NE * ph->dim * 9);
        if (mesh->extraction) {
            const size_t nn = (size_t)ph->n * ph->n;
            for (int64_t e = 0; e < ph->NE; e++)
                if (!factor_extraction(ph->dim, mesh->extraction + e * nn, ph->ext1d.data() + e * ph->dim * 9))
                    THROW(IFE_ERR_UNSUPPORTED, "extraction operator of element " + std::to_string(e) +
                                                   " is not a Kronecker product of 1D operators");
        } else {
            std::copy(mesh->extraction_1d, mesh->extraction_1d + ph->ext1d.size(), ph->ext1d.begin());
        }
        for (int i = 0; i < dofs->n_fields; i++) {
            if (!dofs->names[i]) THROW(IFE_ERR_INVALID, "null dof name");
            ph->fname.push_back(dofs->names[i]);
            if (dofs->steps[i] < 0) THROW(IFE_ERR_INVALID, "negative dof step");
            ph->fstep.push_back(dofs->steps[i]);
            ph->nsteps = std::max(ph->nsteps, dofs->steps[i] + 1);
        }
        ph->fused.assign(ph->fname.size(), 0);
        *out = ph.release();
    });
}

void ife_physics_destroy(ife_physics* ph) {
    if (!ph) return;
    for (auto& st : ph->steps) krylov_destroy(st.kry);
    delete ph;
}

int ife_physics_add_model(ife_physics* ph, const char* type_name, const char* name, int32_t elem_group,
                          const ife_material* mat, const ife_phasefield_opts* pf, const ife_model_opts* opts,
                          int32_t* model_index) {
    return guard([&] {
        if (!ph || !type_name || !mat) THROW(IFE_ERR_INVALID, "null argument");
        if (ph->planned) THROW(IFE_ERR_STATE, "models must be added before setup");
        if (elem_group != 0) THROW(IFE_ERR_UNSUPPORTED, "only element group 0 (the whole mesh) is supported");
        std::string tn(type_name);
        auto m = std::make_unique<Model>();
        m->name = name ? name : tn;
        m->mat = *mat;
        auto need = [&](const char* f) {
            int i = ph->field_index(f);
            if (i < 0) THROW(IFE_ERR_INVALID, std::string("dof type ") + f + " not in DofNames");
            ph->fused[i] = 1;  // AddDofs on the group's unique nodes
            return i;
        };
        // multiplier DOFs of the Lagrange irreversibility methods on the phase nodes, in the
        // phase step (PhaseField.cpp:96-101,270-280; CossPhaseField.cpp:101-106,296-306)
        auto add_multipliers = [&](Model* md) {
            const int irr = md->opt.Irreversible;
            if (irr < IFE_IRR_NONE || irr > IFE_IRR_AUGLAGR3) THROW(IFE_ERR_INVALID, "unknown irreversibility method");
            if (irr < IFE_IRR_LAGRANGE) return;
            md->field_L = need("lMult_pf");
            if (ph->fstep[md->field_L] != md->Step_phase)
                THROW(IFE_ERR_INVALID, "the phase field requires phase and lMult_pf to be in the same solver step");
            if (irr == IFE_IRR_AUGLAGR2 || irr == IFE_IRR_AUGLAGR3) {
                md->field_L2 = need("lMult_pf2");
                if (ph->fstep[md->field_L2] != md->Step_phase)
                    THROW(IFE_ERR_INVALID, "the phase field requires phase and lMult_pf2 to be in the same solver step");
            }
        };
        if (tn == "Fracture/PhaseField") {
            if (ph->pf_model()) THROW(IFE_ERR_UNSUPPORTED, "one Fracture/PhaseField model per physics");
            if (!pf || !opts) THROW(IFE_ERR_INVALID, "phase-field options required");
            m->type = MT_PHASEFIELD;
            m->pf = *pf;
            m->opt = *opts;
            const char* un[3] = {"ux", "uy", "uz"};
            m->nfu = ph->dim;
            for (int d = 0; d < ph->dim; d++) m->field_u[d] = need(un[d]);
            for (int d = 1; d < ph->dim; d++)
                if (ph->fstep[m->field_u[d]] != ph->fstep[m->field_u[0]])
                    THROW(IFE_ERR_INVALID, "Fracture/PhaseField requires ux and uy to be in the same solver step");
            m->Step_u = ph->fstep[m->field_u[0]];
            m->field_phase = need("phase");
            m->Step_phase = ph->fstep[m->field_phase];
            m->HasT = ph->field_index("T") >= 0;  // physics->dofspace->hasDofType("T")
            if (m->HasT) { m->field_T = need("T"); m->Step_T = ph->fstep[m->field_T]; }
            const int r = mat->Rheology;
            // SolidMaterial::UpdatePlasticStrains has no dVEP case and throws (SolidMaterial.cpp:188-191)
            if (!(r == IFE_RHEO_LE || r == IFE_RHEO_VE || r == IFE_RHEO_DVE || r == IFE_RHEO_VE_EXPL || r == IFE_RHEO_DVE_EXPL ||
                  r == IFE_RHEO_VEP))
                THROW(IFE_ERR_UNSUPPORTED, "rheology not available for Fracture/PhaseField (LE, VE, dVE, VE_expl, dVE_expl, VEP are)");
            m->HistSize = (r == IFE_RHEO_LE) ? 0 : (r == IFE_RHEO_VEP ? 4 : 1);  // SolidMaterial.cpp:26-52
            if (mat->Split < IFE_SPLIT_NONE || mat->Split > IFE_SPLIT_STARCONVEX)
                THROW(IFE_ERR_INVALID, "unknown energy split");
            // the reference throws for this combination (PhaseFieldMaterial.cpp:150-152)
            if (mat->Split == IFE_SPLIT_SPECSTRESS && pf->ConsistentEnergy)
                THROW(IFE_ERR_UNSUPPORTED, "SpecStress split with ConsistentEnergy is not defined by the reference");
            int dfn = pf->DistributionFunction;
            if (dfn < IFE_DIST_LINEAR || dfn > IFE_DIST_CUSTOM) THROW(IFE_ERR_INVALID, "unknown distribution function");
            if (dfn == IFE_DIST_CUSTOM && pf->Coefs[0] == 0.0)
                THROW(IFE_ERR_INVALID, "Custom distribution: Coefs[0] (w) must be nonzero");
            add_multipliers(m.get());
            int dams[] = {pf->DamFunction, pf->DensityDamage, pf->GravityDamage, pf->ViscoDriving};
            for (int d : dams)
                if (d < IFE_DAM_NONE || d > IFE_DAM_PORDER) THROW(IFE_ERR_INVALID, "unknown damage function");
        } else if (tn == "Fracture/CossPhaseField") {
            // CossPhaseField::Setup (CossPhaseField.cpp:168-280)
            if (ph->pf_model()) THROW(IFE_ERR_UNSUPPORTED, "one phase-field model per physics");
            if (!pf || !opts) THROW(IFE_ERR_INVALID, "phase-field options required");
            m->type = MT_PHASEFIELD;
            m->coss = true;
            m->pf = *pf;
            m->opt = *opts;
            m->opt.LumpedPFCapacity = 0;  // CossPhaseField has no lumped capacity / driving forms
            m->opt.LumpedDriving = 0;
            // u-step fields: ux uy uz ox oy oz (3D, :172-183), ux uy oz (2D, CossPhaseField.h:15)
            const char* un3[6] = {"ux", "uy", "uz", "ox", "oy", "oz"};
            const char* un2[3] = {"ux", "uy", "oz"};
            m->nfu = ph->dim == 3 ? 6 : 3;
            for (int d = 0; d < m->nfu; d++) m->field_u[d] = need(ph->dim == 3 ? un3[d] : un2[d]);
            for (int d = 1; d < m->nfu; d++)
                if (ph->fstep[m->field_u[d]] != ph->fstep[m->field_u[0]])
                    THROW(IFE_ERR_INVALID, "Fracture/CossPhaseField requires the u and w fields in one solver step");
            m->Step_u = ph->fstep[m->field_u[0]];
            m->field_phase = need("phase");
            m->Step_phase = ph->fstep[m->field_phase];
            m->HasT = ph->field_index("T") >= 0;
            if (m->HasT) { m->field_T = need("T"); m->Step_T = ph->fstep[m->field_T]; }
            const int r = mat->Rheology;
            // CosseratMaterial::UpdatePlasticStrains: LE, VEP, dVEP (CosseratMaterial.cpp:451-506)
            if (!(r == IFE_RHEO_LE || r == IFE_RHEO_VEP || r == IFE_RHEO_DVEP))
                THROW(IFE_ERR_UNSUPPORTED, "rheology not defined for a Cosserat material (LE, VEP, dVEP are)");
            m->HistSize = r == IFE_RHEO_LE ? 0 : 4;
            // CosseratMaterial::EnergySplit (CosseratMaterial.cpp:109-152) throws for the others
            if (!(mat->Split == IFE_SPLIT_NONE || mat->Split == IFE_SPLIT_SPECSTRESS ||
                  mat->Split == IFE_SPLIT_DRUCKERPRAGER || mat->Split == IFE_SPLIT_STARCONVEX))
                THROW(IFE_ERR_UNSUPPORTED, "Cosserat energy split: NoSplit, SpecStress, DruckerPrager or StarConvex");
            int dfn = pf->DistributionFunction;
            if (dfn < IFE_DIST_LINEAR || dfn > IFE_DIST_CUSTOM) THROW(IFE_ERR_INVALID, "unknown distribution function");
            if (dfn == IFE_DIST_CUSTOM && pf->Coefs[0] == 0.0)
                THROW(IFE_ERR_INVALID, "Custom distribution: Coefs[0] (w) must be nonzero");
            add_multipliers(m.get());
            if (mat->l_coss <= 0.0) THROW(IFE_ERR_INVALID, "l_coss must be positive");
        } else if (tn == "Thermal/ThermalModel") {
            if (ph->th_model()) THROW(IFE_ERR_UNSUPPORTED, "one Thermal/ThermalModel per physics");
            m->type = MT_THERMAL;
            m->field_T = need("T");
            m->Step_T = ph->fstep[m->field_T];
        } else {
            THROW(IFE_ERR_NOT_FOUND, "Model " + tn + " has not been defined");  // ModelRegister.cpp:72-75
        }
        ph->models.push_back(m.release());
        if (model_index) *model_index = (int32_t)ph->models.size() - 1;
    });
}

int ife_physics_plan(ife_physics* ph) {
    return guard([&] {
        if (!ph) THROW(IFE_ERR_INVALID, "null physics");
        plan(ph);
    });
}

int ife_physics_setup(ife_physics* ph) {
    return guard([&] {
        if (!ph) THROW(IFE_ERR_INVALID, "null physics");
        setup_device(ph);
    });
}

int ife_physics_num_steps(const ife_physics* ph, int32_t* n_steps) {
    return guard([&] {
        if (!ph || !n_steps) THROW(IFE_ERR_INVALID, "null argument");
        *n_steps = ph->nsteps;
    });
}

int ife_physics_dof_numbers(const ife_physics* ph, int32_t field, int64_t* out) {
    return guard([&] {
        if (!ph || !out) THROW(IFE_ERR_INVALID, "null argument");
        if (!ph->planned) THROW(IFE_ERR_STATE, "plan/setup first");
        if (field < 0 || field >= (int)ph->fname.size()) THROW(IFE_ERR_INVALID, "field out of range");
        int fp = ph->fpos[field];
        for (int64_t v = 0; v < ph->NN; v++) {
            int32_t r = ph->rank[v];
            out[v] = (fp < 0 || r < 0) ? -1 : (int64_t)fp * ph->Ng + r;
        }
    });
}

int ife_physics_pattern(const ife_physics* ph, int32_t step, int64_t* n_rows, int64_t* nnz, const int64_t** row_ptr_host,
                        const int32_t** col_host, const int64_t** row_ptr_dev, const int32_t** col_dev) {
    return guard([&] {
        if (!ph) THROW(IFE_ERR_INVALID, "null physics");
        if (!ph->planned) THROW(IFE_ERR_STATE, "plan/setup first");
        if (step < 0 || step >= ph->nsteps) THROW(IFE_ERR_INVALID, "step out of range");
        const Step& st = ph->steps[step];
        if (n_rows) *n_rows = st.nrows;
        if (nnz) *nnz = (int64_t)st.col.size();
        if (row_ptr_host) *row_ptr_host = st.row_ptr.data();
        if (col_host) *col_host = st.col.data();
        if (row_ptr_dev) *row_ptr_dev = st.d_row_ptr.p;
        if (col_dev) *col_dev = st.d_col.p;
    });
}

int ife_physics_coloring(const ife_physics* ph, int32_t* n_colors, const int64_t** color_ptr, const int32_t** color_elems) {
    return guard([&] {
        if (!ph) THROW(IFE_ERR_INVALID, "null physics");
        if (!ph->planned) THROW(IFE_ERR_STATE, "plan/setup first");
        if (n_colors) *n_colors = (int32_t)ph->color_ptr.size() - 1;
        if (color_ptr) *color_ptr = ph->color_ptr.data();
        if (color_elems) *color_elems = ph->color_elems.data();
    });
}

int ife_physics_state(ife_physics* ph, int32_t step, int32_t kind, double** dev_ptr) {
    return guard([&] {
        require_ready(ph);
        if (step < 0 || step >= ph->nsteps || kind < 0 || kind > IFE_DDSTATE_OLD || !dev_ptr) THROW(IFE_ERR_INVALID, "bad argument");
        *dev_ptr = ph->steps[step].x[kind].p;
    });
}

int ife_physics_set_time(ife_physics* ph, double t, double dt, double du_dt, double ddu_dt) {
    return guard([&] {
        if (!ph) THROW(IFE_ERR_INVALID, "null physics");
        if (!(dt > 0)) THROW(IFE_ERR_INVALID, "dt must be positive");
        ph->t = t; ph->dt = dt; ph->du_dt = du_dt; ph->ddu_dt = ddu_dt;
    });
}

int ife_physics_set_precision(ife_physics* ph, int32_t precision) {
    return guard([&] {
        if (!ph) THROW(IFE_ERR_INVALID, "null physics");
        if (precision != IFE_PREC_FP64 && precision != IFE_PREC_MIXED) THROW(IFE_ERR_INVALID, "unknown precision");
        ph->precision = precision;
    });
}

int ife_physics_assemble(ife_physics* ph, int32_t step, double* K_vals, double* f, void* stream) {
    return guard([&] {
        double* K[1] = {K_vals};
        double* F[1] = {f};
        run(ph, false, step, K, F, (hipStream_t)stream);
    });
}

int ife_physics_assemble_all(ife_physics* ph, double* const* K_vals, double* const* f, void* stream) {
    return guard([&] { run(ph, true, 0, K_vals, f, (hipStream_t)stream); });
}

int ife_physics_adjoint(ife_physics* ph, const double* const* lambda, double* g_Gc, double* g_A, double* g_n,
                        void* stream) {
    return guard([&] { run_adjoint(ph, lambda, g_Gc, g_A, g_n, (hipStream_t)stream); });
}

int ife_physics_spmv(ife_physics* ph, int32_t step, const double* K_vals, const double* x, double* y, void* stream) {
    return guard([&] {
        require_ready(ph);
        if (step < 0 || step >= ph->nsteps) THROW(IFE_ERR_INVALID, "step out of range");
        if (!x || !y) THROW(IFE_ERR_INVALID, "null argument");
        Step& st = ph->steps[step];
        if (!st.has_matrix) THROW(IFE_ERR_STATE, "step has no matrix");
        const double* K = K_vals ? K_vals : st.K.p;
        HIPCHK(launch_spmv(st.d_row_ptr.p, st.d_col.p, K, x, y, st.nrows, (hipStream_t)stream));
    });
}

int ife_physics_set_constraints(ife_physics* ph, int32_t step, int64_t n, const int64_t* dofs, const double* values) {
    return guard([&] {
        require_ready(ph);
        if (step < 0 || step >= ph->nsteps) THROW(IFE_ERR_INVALID, "step out of range");
        if (n < 0 || (n > 0 && (!dofs || !values))) THROW(IFE_ERR_INVALID, "bad constraint arrays");
        Step& st = ph->steps[step];
        if (!st.has_matrix) THROW(IFE_ERR_STATE, "step has no matrix");
        // Constrainer::AddConstraint (Constrainer.cpp:45-54): one value per dof, map order
        std::map<int64_t, double> con;
        for (int64_t i = 0; i < n; i++) {
            const int64_t d = dofs[i];
            if (d < 0 || d >= st.nrows) THROW(IFE_ERR_INVALID, "constrained dof out of range: " + std::to_string(d));
            auto it = con.find(d);
            if (it != con.end()) {
                if (it->second != values[i])
                    THROW(IFE_ERR_INVALID, "Degree of freedom " + std::to_string(d) + " is doubly constrained with different values");
            } else {
                con.emplace(d, values[i]);
            }
        }
        const int64_t N = st.nrows, nc = (int64_t)con.size(), nfree = N - nc;
        std::vector<int64_t> cd, fr;
        std::vector<double> cv;
        std::vector<int64_t> fmap(N, -1);
        cd.reserve(nc); cv.reserve(nc); fr.reserve(nfree);
        {
            auto it = con.begin();
            for (int64_t d = 0; d < N; d++) {
                if (it != con.end() && it->first == d) { cd.push_back(d); cv.push_back(it->second); ++it; }
                else { fmap[d] = (int64_t)fr.size(); fr.push_back(d); }
            }
        }
        // K_con = U^T K U: the free rows, their free columns renumbered (monotone, so sorted)
        // and per free row the entries in constrained columns (the K C c part of f_con)
        std::vector<int64_t> crp(nfree + 1, 0), cfp(nfree + 1, 0);
        for (int64_t k = 0; k < nfree; k++) {
            const int64_t r = fr[k];
            int64_t c = 0;
            for (int64_t q = st.row_ptr[r]; q < st.row_ptr[r + 1]; q++) c += fmap[st.col[q]] >= 0;
            crp[k + 1] = crp[k] + c;
            cfp[k + 1] = cfp[k] + (st.row_ptr[r + 1] - st.row_ptr[r] - c);
        }
        std::vector<int32_t> cfr(cfp[nfree]);
        std::vector<int32_t> ccol(crp[nfree]), rel(crp[nfree]);
        parallel_for(nfree, [&](int64_t a, int64_t b) {
            for (int64_t k = a; k < b; k++) {
                const int64_t r = fr[k];
                int64_t o = crp[k];
                int64_t oc = cfp[k];
                for (int64_t q = st.row_ptr[r]; q < st.row_ptr[r + 1]; q++) {
                    const int64_t m = fmap[st.col[q]];
                    if (m < 0) {
                        cfr[oc++] = (int32_t)(q - st.row_ptr[r]);
                        continue;
                    }
                    ccol[o] = (int32_t)m;
                    rel[o] = (int32_t)(q - st.row_ptr[r]);
                    o++;
                }
            }
        });
        st.has_cons = false;
        st.nfree = nfree;
        st.c_row_ptr = std::move(crp);
        st.c_col = std::move(ccol);
        st.d_cdofs.upload(cd.data(), cd.size());
        st.d_cvals.upload(cv.data(), cv.size());
        st.d_free.upload(fr.data(), fr.size());
        st.d_crp.upload(st.c_row_ptr.data(), st.c_row_ptr.size());
        st.d_ccol.upload(st.c_col.data(), st.c_col.size());
        st.d_rel.upload(rel.data(), rel.size());
        st.d_cfp.upload(cfp.data(), cfp.size());
        st.d_cfr.upload(cfr.data(), cfr.size());
        std::vector<double> zero(N, 0.0);  // g is zero at the free dofs for good
        st.d_g.upload(zero.data(), zero.size());
        st.d_kcon.alloc(st.c_col.size());
        st.d_fcon.alloc(nfree);
        st.d_part.alloc((size_t)3 * norms_partials());
        st.d_norms.alloc(3);
        st.d_ls.alloc(1);
        st.has_cons = true;
    });
}

namespace {
// Physics::Constraint (f_only 0) or Physics::ConstraintFOnly (f_only 1) of a step
void constrain_step(ife_physics* ph, int32_t step, int f_only, const int* gate, hipStream_t s) {
        require_ready(ph);
        if (step < 0 || step >= ph->nsteps) THROW(IFE_ERR_INVALID, "step out of range");
        Step& st = ph->steps[step];
        if (!st.has_cons) THROW(IFE_ERR_STATE, "no constraints set for this step (ife_physics_set_constraints)");
        ConsArgs a{};
        a.f_only = f_only;
        a.gate = gate;
        a.n = st.nrows;
        a.nc = (int64_t)st.d_cdofs.n;
        a.nfree = st.nfree;
        a.krp = st.d_row_ptr.p;
        a.kcol = st.d_col.p;
        a.Kv = st.K.p;
        a.f = st.f.p;
        a.x = st.x[IFE_STATE].p;
        a.cdofs = st.d_cdofs.p;
        a.cvals = st.d_cvals.p;
        a.free = st.d_free.p;
        a.crp = st.d_crp.p;
        a.rel = st.d_rel.p;
        a.cfp = st.d_cfp.p;
        a.cfr = st.d_cfr.p;
        a.g = st.d_g.p;
        a.Kc = st.d_kcon.p;
        a.fcon = st.d_fcon.p;
        HIPCHK(launch_constrain(a, s));
}
}  // namespace

int ife_physics_constrain(ife_physics* ph, int32_t step, void* stream) {
    return guard([&] { constrain_step(ph, step, 0, nullptr, (hipStream_t)stream); });
}

int ife_physics_constrain_f(ife_physics* ph, int32_t step, void* stream) {
    return guard([&] { constrain_step(ph, step, 1, nullptr, (hipStream_t)stream); });
}

int ife_physics_constrained(const ife_physics* ph, int32_t step, int64_t* n_free, int64_t* nnz,
                            const int64_t** row_ptr_host, const int32_t** col_host, const int64_t** row_ptr_dev,
                            const int32_t** col_dev, double** K_con, double** f_con) {
    return guard([&] {
        require_ready(ph);
        if (step < 0 || step >= ph->nsteps) THROW(IFE_ERR_INVALID, "step out of range");
        const Step& st = ph->steps[step];
        if (!st.has_cons) THROW(IFE_ERR_STATE, "no constraints set for this step (ife_physics_set_constraints)");
        if (n_free) *n_free = st.nfree;
        if (nnz) *nnz = (int64_t)st.c_col.size();
        if (row_ptr_host) *row_ptr_host = st.c_row_ptr.data();
        if (col_host) *col_host = st.c_col.data();
        if (row_ptr_dev) *row_ptr_dev = st.d_crp.p;
        if (col_dev) *col_dev = st.d_ccol.p;
        if (K_con) *K_con = st.d_kcon.p;
        if (f_con) *f_con = st.d_fcon.p;
    });
}

int ife_physics_residual_norms_async(ife_physics* ph, int32_t step, const double* dx, double* out_dev, void* stream) {
    return guard([&] {
        require_ready(ph);
        if (step < 0 || step >= ph->nsteps) THROW(IFE_ERR_INVALID, "step out of range");
        if (!dx || !out_dev) THROW(IFE_ERR_INVALID, "null argument");
        Step& st = ph->steps[step];
        if (!st.has_cons) THROW(IFE_ERR_STATE, "no constraints set for this step (ife_physics_set_constraints)");
        HIPCHK(launch_norms(dx, st.d_fcon.p, st.nfree, st.d_part.p, out_dev, (hipStream_t)stream));
    });
}

int ife_physics_residual_norms(ife_physics* ph, int32_t step, const double* dx, void* stream, double out[3]) {
    return guard([&] {
        require_ready(ph);
        if (step < 0 || step >= ph->nsteps) THROW(IFE_ERR_INVALID, "step out of range");
        if (!dx || !out) THROW(IFE_ERR_INVALID, "null argument");
        Step& st = ph->steps[step];
        if (!st.has_cons) THROW(IFE_ERR_STATE, "no constraints set for this step (ife_physics_set_constraints)");
        hipStream_t s = (hipStream_t)stream;
        HIPCHK(launch_norms(dx, st.d_fcon.p, st.nfree, st.d_part.p, st.d_norms.p, s));
        HIPCHK(hipMemcpyAsync(out, st.d_norms.p, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    });
}

int ife_physics_solve(ife_physics* ph, int32_t step, const ife_solve_opts* opts, double* dx_dev, void* stream,
                      ife_solve_info* info) {
    return guard([&] {
        require_ready(ph);
        if (step < 0 || step >= ph->nsteps) THROW(IFE_ERR_INVALID, "step out of range");
        if (!dx_dev) THROW(IFE_ERR_INVALID, "null dx");
        Step& st = ph->steps[step];
        KrylovOpts o{1.0e-9, 1.0e-15, 100, 10000, IFE_PC_ILU0};  // GMRESLinSolver.cpp:23-24
        if (opts) {
            o = KrylovOpts{opts->rtol, opts->atol, opts->restart, opts->max_iterations, opts->preconditioner};
            if (!(o.rtol >= 0) || !(o.atol >= 0)) THROW(IFE_ERR_INVALID, "tolerances must be non-negative");
            if (o.restart < 1 || o.restart > 1000) THROW(IFE_ERR_INVALID, "restart must be in [1, 1000]");
            if (o.max_it < 1) THROW(IFE_ERR_INVALID, "max_iterations must be positive");
            if (o.pc < IFE_PC_NONE || o.pc > IFE_PC_ILU0) THROW(IFE_ERR_INVALID, "unknown preconditioner");
        }
        KrylovSystem sys;
        if (st.has_cons) {
            sys = KrylovSystem{st.nfree, (int64_t)st.c_col.size(), st.d_crp.p, st.d_ccol.p, st.d_kcon.p, st.d_fcon.p, 1.0};
        } else {
            sys = KrylovSystem{st.nrows, (int64_t)st.col.size(), st.d_row_ptr.p, st.d_col.p, st.K.p, st.f.p, -1.0};
        }
        if (!st.kry) st.kry = krylov_create();
        KrylovResult r;
        std::string err;
        const int rc = krylov_solve(st.kry, sys, o, dx_dev, (hipStream_t)stream, &r, &err);
        if (rc != IFE_OK) THROW(rc, "ife_physics_solve: " + err);
        if (info) *info = ife_solve_info{r.iterations, r.restarts, r.converged, r.res_norm, r.rhs_norm};
    });
}

int ife_residual_norms_noalloc(const double* dx, const double* f_con, int64_t n, double* out_dev, void* stream) {
    return guard([&] {
        if (!dx || !f_con || !out_dev) THROW(IFE_ERR_INVALID, "null argument");
        if (n <= 0 || n > IFE_NOALLOC_NORMS_MAX_N)
            THROW(IFE_ERR_INVALID, "ife_residual_norms_noalloc: n must be in [1, 2^53] (got " + std::to_string(n) + ")");
        HIPCHK(launch_norms_noalloc(dx, f_con, n, out_dev, (hipStream_t)stream));
    });
}

int ife_physics_set_time_scheme(ife_physics* ph, int32_t scheme, double c0, double c1) {
    return guard([&] {
        if (!ph) THROW(IFE_ERR_INVALID, "null physics");
        if (scheme < IFE_TS_NEWMARK || scheme > IFE_TS_BDF2) THROW(IFE_ERR_INVALID, "Time Discretisation type not defined");
        if (scheme == IFE_TS_NEWMARK && !(c0 != 0.0)) THROW(IFE_ERR_INVALID, "Newmark beta must be non-zero");
        if (scheme == IFE_TS_ALPHA && !(c0 != 0.0)) THROW(IFE_ERR_INVALID, "alpha must be non-zero");
        ph->ts = scheme;
        ph->ts_c0 = scheme == IFE_TS_EULER ? 1.0 : c0;
        ph->ts_c1 = c1;
    });
}

int ife_physics_set_dt(ife_physics* ph, double t, double dt) {
    return guard([&] {
        if (!ph) THROW(IFE_ERR_INVALID, "null physics");
        if (!(dt > 0)) THROW(IFE_ERR_INVALID, "dt must be positive");
        double du, ddu;
        if (ph->ts == IFE_TS_NEWMARK) { du = ph->ts_c1 / (dt * ph->ts_c0); ddu = 1.0 / (dt * dt * ph->ts_c0); }
        else if (ph->ts == IFE_TS_BDF2) { du = 1.0 / dt; ddu = 1.0 / (dt * dt); }
        else { du = 1.0 / (dt * ph->ts_c0); ddu = 0.0; }
        ph->t = t; ph->dt = dt; ph->du_dt = du; ph->ddu_dt = ddu;
    });
}

static void vel_acc_all(ife_physics* ph, hipStream_t s, const int* gate = nullptr) {
    for (int q = 0; q < ph->nsteps; q++) {
        Step& st = ph->steps[q];
        HIPCHK(launch_vel_acc(ph->ts, ph->ts_c0, ph->ts_c1, ph->dt, st.x[IFE_STATE].p, st.x[IFE_STATE_OLD].p,
                              st.x[IFE_DSTATE_OLD].p, st.x[IFE_DDSTATE_OLD].p, st.x[IFE_DSTATE].p, st.x[IFE_DDSTATE].p,
                              st.nrows, s, gate));
    }
}

int ife_physics_update_state(ife_physics* ph, int32_t step, const double* dx, void* stream) {
    return guard([&] {
        require_ready(ph);
        if (step < 0 || step >= ph->nsteps) THROW(IFE_ERR_INVALID, "step out of range");
        if (!dx) THROW(IFE_ERR_INVALID, "null increment");
        Step& st = ph->steps[step];
        hipStream_t s = (hipStream_t)stream;
        if (st.has_cons)
            HIPCHK(launch_update_state(st.x[IFE_STATE].p, dx, st.d_free.p, st.nfree, st.d_cdofs.p, st.d_g.p,
                                       (int64_t)st.d_cdofs.n, s));
        else
            HIPCHK(launch_update_state(st.x[IFE_STATE].p, dx, nullptr, st.nrows, nullptr, nullptr, 0, s));
        vel_acc_all(ph, s);
    });
}

int ife_physics_line_search(ife_physics* ph, int32_t step, double* dx, double lim_lo, double lim_hi, double* eta_dev,
                            void* stream) {
    return guard([&] {
        require_ready(ph);
        if (step < 0 || step >= ph->nsteps) THROW(IFE_ERR_INVALID, "step out of range");
        if (!dx) THROW(IFE_ERR_INVALID, "null increment");
        if (!(lim_lo > 0.0) || !(lim_hi >= lim_lo)) THROW(IFE_ERR_INVALID, "line-search limits must satisfy 0 < lo <= hi");
        Step& st = ph->steps[step];
        if (!st.has_cons) THROW(IFE_ERR_STATE, "the line search works on the constrained system (ife_physics_set_constraints)");
        hipStream_t s = (hipStream_t)stream;
        LSDev* ls = st.d_ls.p;
        double* x = st.x[IFE_STATE].p;
        const int64_t nc = (int64_t)st.d_cdofs.n;
        HIPCHK(launch_ls_init(ls, st.d_fcon.p, dx, st.nfree, s));
        for (int it = 0; it < LS_MAX; it++) {
            const int* gate = &ls->run[it];
            const double* L = &ls->L[it];
            // UpdateState(L dx), AssembleFOnly, ConstraintFOnly, e = f_con . (L dx) (:320-327)
            HIPCHK(launch_update_state(x, dx, st.d_free.p, st.nfree, st.d_cdofs.p, st.d_g.p, nc, s, L, 1.0, gate));
            vel_acc_all(ph, s, gate);
            run(ph, false, step, nullptr, nullptr, s, gate);
            constrain_step(ph, step, 1, gate, s);
            HIPCHK(launch_ls_trial_end(ls, it, st.d_fcon.p, dx, st.nfree, lim_lo, lim_hi, s));
            // UpdateState(-L dx) (:354-356)
            HIPCHK(launch_update_state(x, dx, st.d_free.p, st.nfree, st.d_cdofs.p, st.d_g.p, nc, s, L, -1.0, gate));
            vel_acc_all(ph, s, gate);
        }
        // dx *= Lcorr, UpdateState(dx) (:379-380)
        HIPCHK(launch_ls_finish(ls, dx, st.nfree, eta_dev, s));
        HIPCHK(launch_update_state(x, dx, st.d_free.p, st.nfree, st.d_cdofs.p, st.d_g.p, nc, s));
        vel_acc_all(ph, s);
        // SolveStep_Standard: Assemble + Constraint when eta < 1 - 1e-6 (NonLinSolver.cpp:170-175)
        run(ph, false, step, nullptr, nullptr, s, &ls->reassemble);
        constrain_step(ph, step, 0, &ls->reassemble, s);
    });
}

int ife_physics_line_search_info(ife_physics* ph, int32_t step, void* stream, ife_line_search_info* out) {
    return guard([&] {
        require_ready(ph);
        if (step < 0 || step >= ph->nsteps) THROW(IFE_ERR_INVALID, "step out of range");
        if (!out) THROW(IFE_ERR_INVALID, "null argument");
        Step& st = ph->steps[step];
        if (!st.has_cons) THROW(IFE_ERR_STATE, "no constraints set for this step (ife_physics_set_constraints)");
        LSDev h;
        HIPCHK(hipMemcpyAsync(&h, st.d_ls.p, sizeof(LSDev), hipMemcpyDeviceToHost, (hipStream_t)stream));
        HIPCHK(hipStreamSynchronize((hipStream_t)stream));
        out->trials = h.trials;
        out->eta = h.eta;
        out->e0 = h.e0;
        out->e_min = h.eMin;
        out->reassembled = h.reassemble;
        for (int i = 0; i < 10; i++) {
            out->L[i] = h.L[i];
            out->e[i] = h.e[i];
        }
    });
}

int ife_physics_commit_state(ife_physics* ph, void* stream) {
    return guard([&] {
        require_ready(ph);
        hipStream_t s = (hipStream_t)stream;
        vel_acc_all(ph, s);
        for (int q = 0; q < ph->nsteps; q++) {
            Step& st = ph->steps[q];
            HIPCHK(launch_copy(st.x[IFE_STATE_OLD].p, st.x[IFE_STATE].p, st.nrows, s));
            HIPCHK(launch_copy(st.x[IFE_DSTATE_OLD].p, st.x[IFE_DSTATE].p, st.nrows, s));
            HIPCHK(launch_copy(st.x[IFE_DDSTATE_OLD].p, st.x[IFE_DDSTATE].p, st.nrows, s));
        }
    });
}

int ife_physics_reset_state(ife_physics* ph, void* stream) {
    return guard([&] {
        require_ready(ph);
        hipStream_t s = (hipStream_t)stream;
        for (int q = 0; q < ph->nsteps; q++) {
            Step& st = ph->steps[q];
            HIPCHK(launch_copy(st.x[IFE_STATE].p, st.x[IFE_STATE_OLD].p, st.nrows, s));
            HIPCHK(launch_copy(st.x[IFE_DSTATE].p, st.x[IFE_DSTATE_OLD].p, st.nrows, s));
            HIPCHK(launch_copy(st.x[IFE_DDSTATE].p, st.x[IFE_DDSTATE_OLD].p, st.nrows, s));
        }
    });
}

int ife_physics_Kf(ife_physics* ph, int32_t step, double** K_vals, double** f) {
    return guard([&] {
        require_ready(ph);
        if (step < 0 || step >= ph->nsteps) THROW(IFE_ERR_INVALID, "step out of range");
        if (K_vals) *K_vals = ph->steps[step].K.p;
        if (f) *f = ph->steps[step].f.p;
    });
}

int ife_physics_status(ife_physics* ph, void* stream, ife_assembly_status* out) {
    return guard([&] {
        require_ready(ph);
        if (!out) THROW(IFE_ERR_INVALID, "null argument");
        hipStream_t s = (hipStream_t)stream;
        Status st;
        HIPCHK(hipMemcpyAsync(&st, ph->d_status.p, sizeof(Status), hipMemcpyDeviceToHost, s));
        std::vector<double> lc(ph->NE);
        HIPCHK(hipMemcpyAsync(lc.data(), ph->d_lcrack.p, ph->NE * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        out->elements_assembled = (int64_t)st.elements;
        out->return_map_calls = (int64_t)st.rm_calls;
        out->return_map_iterations = (int64_t)st.rm_iters;
        out->return_map_maxit_hits = (int64_t)st.rm_maxit;
        out->nonfinite_seen = (int32_t)st.nonfinite;
        double s_ = 0;
        if (ph->phase_assembled)
            for (int64_t e = 0; e < ph->NE; e++) s_ += lc[e];
        out->LCrack = s_;
    });
}

int ife_physics_set_profiling(ife_physics* ph, int32_t enable) {
    return guard([&] {
        if (!ph) THROW(IFE_ERR_INVALID, "null physics");
        ph->profiling = enable != 0;
    });
}

int ife_physics_timings(ife_physics* ph, ife_timings* out) {
    return guard([&] {
        if (!ph || !out) THROW(IFE_ERR_INVALID, "null argument");
        out->launches = 0; out->kernel_ms_total = 0; out->kernel_ms_max = 0; out->elements = 0;
        for (auto& r : ph->recs) {
            HIPCHK(hipEventSynchronize(r.b));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
            out->launches++;
            out->kernel_ms_total += ms;
            out->kernel_ms_max = std::max(out->kernel_ms_max, (double)ms);
            out->elements += r.elems;
            ph->event_pool.push_back(r.a);
            ph->event_pool.push_back(r.b);
        }
        ph->recs.clear();
        out->material_ms_total = 0;
        out->material_launches = 0;
        for (auto& r : ph->mat_recs) {
            HIPCHK(hipEventSynchronize(r.b));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
            out->material_launches++;
            out->material_ms_total += ms;
            ph->event_pool.push_back(r.a);
            ph->event_pool.push_back(r.b);
        }
        ph->mat_recs.clear();
        out->gather_ms_total = 0;
        for (auto& r : ph->gat_recs) {
            HIPCHK(hipEventSynchronize(r.b));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
            out->gather_ms_total += ms;
            ph->event_pool.push_back(r.a);
            ph->event_pool.push_back(r.b);
        }
        ph->gat_recs.clear();
    });
}

int ife_model_history(ife_physics* ph, int32_t model, const char* name, double** dev_ptr, int64_t* n_components,
                      int64_t* n_per_component) {
    return guard([&] {
        require_ready(ph);
        if (model < 0 || model >= (int)ph->models.size() || !name) THROW(IFE_ERR_INVALID, "bad model index");
        Model* m = ph->models[model];
        if (m->type != MT_PHASEFIELD) THROW(IFE_ERR_NOT_FOUND, "model has no history");
        int64_t nc = 0;
        DevBuf<double>* b = history_buf(m, name, nc);
        if (!b) THROW(IFE_ERR_NOT_FOUND, std::string("unknown history ") + name);
        if (dev_ptr) *dev_ptr = b->p;
        if (n_components) *n_components = nc;
        if (n_per_component) *n_per_component = ph->NE * ph->nip;
    });
}

static int copy_history(ife_physics* ph, int32_t model, bool commit, void* stream) {
    return guard([&] {
        require_ready(ph);
        if (model < 0 || model >= (int)ph->models.size()) THROW(IFE_ERR_INVALID, "bad model index");
        Model* m = ph->models[model];
        if (m->type != MT_PHASEFIELD) return;  // ThermalModel keeps no history
        hipStream_t s = (hipStream_t)stream;
        std::pair<DevBuf<double>*, DevBuf<double>*> pr[] = {
            {&m->E, &m->EOld}, {&m->PS, &m->PSOld}, {&m->VD, &m->VDOld}, {&m->DVD, &m->DVDOld}, {&m->MH, &m->MHOld}};
        for (auto& q : pr) {
            DevBuf<double>* cur = q.first;
            DevBuf<double>* old = q.second;
            if (commit) HIPCHK(launch_copy(old->p, cur->p, (int64_t)cur->n, s));
            else HIPCHK(launch_copy(cur->p, old->p, (int64_t)cur->n, s));
        }
    });
}

int ife_model_set_element_params(ife_physics* ph, int32_t model, const double* Gc_e, const double* A_e) {
    return guard([&] {
        require_ready(ph);
        if (model < 0 || model >= (int)ph->models.size()) THROW(IFE_ERR_INVALID, "bad model index");
        Model* m = ph->models[model];
        if (m->type != MT_PHASEFIELD) THROW(IFE_ERR_INVALID, "per-element parameters belong to a Fracture/PhaseField model");
        if (Gc_e) HIPCHK(hipMemcpy(m->Gc_e.p, Gc_e, ph->NE * sizeof(double), hipMemcpyHostToDevice));
        if (A_e) HIPCHK(hipMemcpy(m->A_e.p, A_e, ph->NE * sizeof(double), hipMemcpyHostToDevice));
        m->use_Gc_e = Gc_e != nullptr;
        m->use_A_e = A_e != nullptr;
    });
}

int ife_model_set_param(ife_physics* ph, int32_t model, const char* name, double value) {
    return guard([&] {
        if (!ph || !name) THROW(IFE_ERR_INVALID, "null argument");
        if (model < 0 || model >= (int)ph->models.size()) THROW(IFE_ERR_INVALID, "bad model index");
        Model* m = ph->models[model];
        std::string nm(name);
        if (nm == "elements_per_wave") {
            if (!(value == 1.0 || value == 4.0 || value == 8.0 || value == 16.0))
                THROW(IFE_ERR_INVALID, "elements_per_wave must be 1, 4, 8 or 16");
            const int v = (int)value;
            if (v > 1 && ph->dim != 2) THROW(IFE_ERR_INVALID, "elements_per_wave above 1 is a 2D element kernel");
            ph->epw2d = v;
            return;
        }
        if (m->type != MT_PHASEFIELD) THROW(IFE_ERR_INVALID, "parameters belong to a Fracture/PhaseField model");
        if (nm == "n") m->n_override = value;
        else if (nm == "rm_abs_tol") m->rm_eabs = value;
        else if (nm == "rm_rel_tol") m->rm_erel = value;
        else if (nm == "rm_max_iterations") m->rm_maxit = (int)value;
        else if (nm == "exact_tangent") m->exact_tangent = value != 0.0;
        else if (nm == "coss_record_chunk") {
            // elements per u-step record chunk: 0 as many as fit, negative: no records (the
            // element blocks added into the CSR colour by colour)
            if (!m->coss) THROW(IFE_ERR_INVALID, "coss_record_chunk belongs to a Fracture/CossPhaseField model");
            ph->crec_chunk = (int64_t)value;
        }
        else THROW(IFE_ERR_NOT_FOUND, "unknown parameter " + nm);
    });
}

int ife_model_commit(ife_physics* ph, int32_t model, int32_t commit_type, void* stream) {
    if (commit_type != 1) return IFE_OK;  // only TIMEDEP_COMMIT_TYPE copies history (PhaseField.cpp:124)
    return copy_history(ph, model, true, stream);
}

int ife_model_reset_step(ife_physics* ph, int32_t model, void* stream) { return copy_history(ph, model, false, stream); }

int ife_batch_create(ife_physics* const* domains, int32_t n_domains, ife_batch** out) {
    return guard([&] {
        if (!out || (!domains && n_domains > 0) || n_domains <= 0) THROW(IFE_ERR_INVALID, "batch: no domains");
        *out = nullptr;
        auto b = std::make_unique<ife_batch>();
        std::vector<int64_t> pre[3];
        for (int k = 0; k < 3; k++) pre[k].push_back(0);
        for (int32_t d = 0; d < n_domains; d++) {
            ife_physics* ph = domains[d];
            if (!ph) THROW(IFE_ERR_INVALID, "batch: null domain");
            require_ready(ph);
            if (d == 0) { b->dim = ph->dim; b->nta = ph->dim == 2 ? 64 : 256; }
            if (ph->dim != b->dim) THROW(IFE_ERR_UNSUPPORTED, "batch: domains need the same dimension");
            b->doms.push_back(ph);
            pre[0].push_back(pre[0].back() + blocks_material(ph->dim, ph->NE));
            pre[1].push_back(pre[1].back() + ph->NE);
            const int npw = gather_nodes_per_wave(ph->dim);
            pre[2].push_back(pre[2].back() + (ph->Ng + npw - 1) / npw);
            b->nodes += ph->Ng;
        }
        for (int k = 0; k < 3; k++) {
            b->bp[k].upload(pre[k].data(), pre[k].size());
            b->tot[k] = pre[k].back();
        }
        if (b->dim == 2) {  // MFMA element kernel: ceil(elements / epw) wavefronts per domain
            const int epws[3] = {4, 8, 16};
            for (int k = 0; k < 3; k++) {
                std::vector<int64_t> pw{0};
                for (auto* ph : b->doms) pw.push_back(pw.back() + (ph->NE + epws[k] - 1) / epws[k]);
                b->bpw[k].upload(pw.data(), pw.size());
                b->totw[k] = pw.back();
            }
        }
        batch_update(b.get());
        *out = b.release();
    });
}

int ife_batch_update(ife_batch* b) {
    return guard([&] {
        if (!b) THROW(IFE_ERR_INVALID, "null batch");
        batch_update(b);
    });
}

int ife_batch_assemble(ife_batch* b, void* stream) {
    return guard([&] {
        if (!b) THROW(IFE_ERR_INVALID, "null batch");
        batch_enqueue(b, (hipStream_t)stream);
    });
}

int ife_batch_capture(ife_batch* b) {
    return guard([&] {
        if (!b) THROW(IFE_ERR_INVALID, "null batch");
        batch_capture(b);
    });
}

int ife_batch_launch(ife_batch* b, void* stream) {
    return guard([&] {
        if (!b) THROW(IFE_ERR_INVALID, "null batch");
        if (!b->exec) batch_capture(b);
        HIPCHK(hipGraphLaunch(b->exec, (hipStream_t)stream));
        for (auto* ph : b->doms) ph->phase_assembled |= ph->pf_model() != nullptr;
    });
}

int ife_batch_tune(ife_batch* b, void* stream, ife_batch_tuning* out) {
    return guard([&] {
        if (!b) THROW(IFE_ERR_INVALID, "null batch");
        batch_tune(b, (hipStream_t)stream, out);
    });
}

int ife_batch_set_shape(ife_batch* b, const ife_batch_tuning* shape) {
    return guard([&] {
        if (!b || !shape) THROW(IFE_ERR_INVALID, "null argument");
        if (shape->grid_material < 0 || shape->grid_assemble < 0 || shape->grid_gather < 0)
            THROW(IFE_ERR_INVALID, "batch: negative grid");
        const int lanes = b->dim == 2 ? 64 : 256;
        if (shape->lanes_assemble != 0 && shape->lanes_assemble != lanes)
            THROW(IFE_ERR_INVALID, "batch: element kernel lanes per element must be 256 (3D) or 64 (2D, one wavefront), or 0");
        const int e = shape->elems_per_wave;
        if (!(e == 0 || e == 1 || e == 4 || e == 8 || e == 16))
            THROW(IFE_ERR_INVALID, "batch: elements per wavefront must be 0, 1, 4, 8 or 16");
        if (e > 1 && !(b->dim == 2 && b->mixed))
            THROW(IFE_ERR_INVALID, "batch: 4, 8 or 16 elements per wavefront (MFMA) is the 2D mixed-precision kernel");
        if (e > 1 && shape->grid_assemble != 0)
            THROW(IFE_ERR_INVALID, "batch: the MFMA element kernel launches one wavefront per element group (grid 0)");
        b->grid[0] = shape->grid_material;
        b->grid[1] = shape->grid_assemble;
        b->grid[2] = shape->grid_gather;
        b->nta = lanes;
        b->epw = e > 1 ? e : 1;
        b->drop_graph();
    });
}

int ife_batch_get_shape(const ife_batch* b, ife_batch_tuning* out) {
    return guard([&] {
        if (!b || !out) THROW(IFE_ERR_INVALID, "null argument");
        std::memset(out, 0, sizeof(*out));
        out->grid_material = b->grid[0];
        out->grid_assemble = b->grid[1];
        out->grid_gather = b->grid[2];
        out->lanes_assemble = b->nta;
        out->elems_per_wave = b->epw;
        out->n_domains = (int32_t)b->doms.size();
        out->blocks_material = b->tot[0];
        out->elements = b->tot[1];
        out->nodes = b->nodes;
    });
}

void ife_batch_destroy(ife_batch* b) { delete b; }

}  // extern "C"
