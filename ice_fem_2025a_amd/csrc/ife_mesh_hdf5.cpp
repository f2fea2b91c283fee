// ife_mesh_hdf5.cpp -- IceFEM's HDF5 mesh format, read and written on the host.
//
// Layout (TestCases/TriAxialCompression/Meshgen/SaveToHDF.m:1-83, written from MATLAB, so the
// HDF5 dimensions are the MATLAB sizes reversed; read by Mesh::LoadFromFile,
// mesh/mesh.cpp:53-303):
//   /nodes                       double [dim][n_nodes]
//   /node_DataTypes              strings, e.g. {"W"};  /node_Data  double [n_data][n_nodes]
//   /nodegroupnames              strings;  /nodegroups/<name>  uint64 [n] (0-based node ids)
//   /elementgroupnames, /elementgrouptypes   strings (e.g. "internal", "NURBS3_Cube2")
//   /elementgroups/<name>        uint64 [nodes_per_elem][n_elems]  (Elems[e][j] = file[j][e],
//                                 elementgroup.cpp:17-34)
//   /elementgroups/<name>_Data   double [n][n][n_elems]  (ElemData(j,k) = file[k][j][e],
//                                 elementgroup.cpp:40-67)
//   /NCores, /Partition/<rank>/* the MPI partition (unused on one rank, mesh.cpp:61-67)
// The HDF5 C library is opened with dlopen at the first call, so the engine itself does not
// depend on it; without it these entry points return IFE_ERR_UNSUPPORTED.
#include <dlfcn.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "ife.h"
#include "ife_internal.h"

namespace {

// ---- the part of the HDF5 1.10+ C API used here (H5public.h, H5Ipublic.h) ---------------
typedef int64_t hid_t;
typedef int herr_t;
typedef int htri_t;
typedef unsigned long long hsize_t;
constexpr hid_t H5P_DEFAULT_ = 0, H5S_ALL_ = 0, H5E_DEFAULT_ = 0;
constexpr unsigned H5F_ACC_RDONLY_ = 0x0000u, H5F_ACC_TRUNC_ = 0x0002u;
constexpr int H5T_STRING_ = 3, H5T_CSET_UTF8_ = 1;
constexpr size_t H5T_VARIABLE_ = (size_t)(-1);

struct H5 {
    void* so = nullptr;
    herr_t (*open)();
    herr_t (*eset_auto2)(hid_t, void*, void*);
    hid_t (*fopen)(const char*, unsigned, hid_t);
    hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t);
    herr_t (*fclose)(hid_t);
    htri_t (*lexists)(hid_t, const char*, hid_t);
    hid_t (*dopen2)(hid_t, const char*, hid_t);
    hid_t (*dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
    hid_t (*dget_space)(hid_t);
    hid_t (*dget_type)(hid_t);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
    herr_t (*dclose)(hid_t);
    herr_t (*dvlen_reclaim)(hid_t, hid_t, hid_t, void*);  // 1.10; H5Treclaim from 1.12
    int (*sget_ndims)(hid_t);
    int (*sget_dims)(hid_t, hsize_t*, hsize_t*);
    hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*);
    herr_t (*sclose)(hid_t);
    int (*tget_class)(hid_t);
    htri_t (*tis_variable_str)(hid_t);
    size_t (*tget_size)(hid_t);
    hid_t (*tcopy)(hid_t);
    herr_t (*tset_size)(hid_t, size_t);
    herr_t (*tset_cset)(hid_t, int);
    int (*tget_cset)(hid_t);
    herr_t (*tclose)(hid_t);
    hid_t (*gcreate2)(hid_t, const char*, hid_t, hid_t, hid_t);
    herr_t (*gclose)(hid_t);
    hid_t *t_double, *t_uint64, *t_uint8, *t_c_s1;
};

std::mutex g_h5_mu;
H5 g_h5;
bool g_h5_ok = false;

template <class F>
void sym(void* so, const char* name, F& out, bool required = true) {
    out = reinterpret_cast<F>(dlsym(so, name));
    if (!out && required) throw std::runtime_error(std::string("libhdf5 lacks ") + name);
}

const H5& h5() {
    std::lock_guard<std::mutex> lk(g_h5_mu);
    if (g_h5_ok) return g_h5;
    std::vector<std::string> cands;
    if (const char* env = std::getenv("IFE_HDF5_LIB")) cands.push_back(env);
    cands.insert(cands.end(), {"libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"});
    void* so = nullptr;
    for (auto& c : cands)
        if ((so = dlopen(c.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
    if (!so) throw std::invalid_argument("libhdf5 not found (set IFE_HDF5_LIB to its path)");
    H5 h;
    h.so = so;
    sym(so, "H5open", h.open);
    sym(so, "H5Eset_auto2", h.eset_auto2);
    sym(so, "H5Fopen", h.fopen);
    sym(so, "H5Fcreate", h.fcreate);
    sym(so, "H5Fclose", h.fclose);
    sym(so, "H5Lexists", h.lexists);
    sym(so, "H5Dopen2", h.dopen2);
    sym(so, "H5Dcreate2", h.dcreate2);
    sym(so, "H5Dget_space", h.dget_space);
    sym(so, "H5Dget_type", h.dget_type);
    sym(so, "H5Dread", h.dread);
    sym(so, "H5Dwrite", h.dwrite);
    sym(so, "H5Dclose", h.dclose);
    sym(so, "H5Dvlen_reclaim", h.dvlen_reclaim, false);
    if (!h.dvlen_reclaim) sym(so, "H5Treclaim", h.dvlen_reclaim);
    sym(so, "H5Sget_simple_extent_ndims", h.sget_ndims);
    sym(so, "H5Sget_simple_extent_dims", h.sget_dims);
    sym(so, "H5Screate_simple", h.screate_simple);
    sym(so, "H5Sclose", h.sclose);
    sym(so, "H5Tget_class", h.tget_class);
    sym(so, "H5Tis_variable_str", h.tis_variable_str);
    sym(so, "H5Tget_size", h.tget_size);
    sym(so, "H5Tcopy", h.tcopy);
    sym(so, "H5Tset_size", h.tset_size);
    sym(so, "H5Tset_cset", h.tset_cset);
    sym(so, "H5Tget_cset", h.tget_cset);
    sym(so, "H5Tclose", h.tclose);
    sym(so, "H5Gcreate2", h.gcreate2);
    sym(so, "H5Gclose", h.gclose);
    sym(so, "H5T_NATIVE_DOUBLE_g", h.t_double);
    sym(so, "H5T_NATIVE_UINT64_g", h.t_uint64);
    sym(so, "H5T_NATIVE_UINT8_g", h.t_uint8);
    sym(so, "H5T_C_S1_g", h.t_c_s1);
    if (h.open() < 0) throw std::runtime_error("H5open failed");
    h.eset_auto2(H5E_DEFAULT_, nullptr, nullptr);  // errors come back as codes, not prints
    g_h5 = h;
    g_h5_ok = true;
    return g_h5;
}

struct Err : std::runtime_error {
    int code;
    Err(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// closes an HDF5 handle at scope exit
struct Hid {
    hid_t id;
    herr_t (*close)(hid_t);
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Hid() { if (id >= 0) close(id); }
    Hid(const Hid&) = delete;
    Hid& operator=(const Hid&) = delete;
};

bool exists(hid_t f, const std::string& path) {
    // H5Lexists on each prefix: a missing intermediate group is "absent", not an error
    size_t pos = 1;
    while (true) {
        size_t nx = path.find('/', pos);
        std::string pre = path.substr(0, nx);
        if (h5().lexists(f, pre.c_str(), H5P_DEFAULT_) <= 0) return false;
        if (nx == std::string::npos) return true;
        pos = nx + 1;
    }
}

std::vector<hsize_t> dims_of(hid_t f, const std::string& path) {
    const H5& h = h5();
    Hid d(h.dopen2(f, path.c_str(), H5P_DEFAULT_), h.dclose);
    if (d.id < 0) throw Err(IFE_ERR_INVALID, "mesh file: cannot open dataset " + path);
    Hid sp(h.dget_space(d.id), h.sclose);
    const int nd = h.sget_ndims(sp.id);
    std::vector<hsize_t> dims(nd > 0 ? nd : 0);
    if (nd > 0) h.sget_dims(sp.id, dims.data(), nullptr);
    return dims;
}

template <class T>
std::vector<T> read_num(hid_t f, const std::string& path, hid_t memtype, std::vector<hsize_t>& dims) {
    const H5& h = h5();
    dims = dims_of(f, path);
    size_t n = 1;
    for (auto d : dims) n *= (size_t)d;
    std::vector<T> out(n);
    Hid d(h.dopen2(f, path.c_str(), H5P_DEFAULT_), h.dclose);
    if (n && h.dread(d.id, memtype, H5S_ALL_, H5S_ALL_, H5P_DEFAULT_, out.data()) < 0)
        throw Err(IFE_ERR_INVALID, "mesh file: cannot read " + path);
    return out;
}

std::vector<std::string> read_strings(hid_t f, const std::string& path) {
    const H5& h = h5();
    Hid d(h.dopen2(f, path.c_str(), H5P_DEFAULT_), h.dclose);
    if (d.id < 0) throw Err(IFE_ERR_INVALID, "mesh file: cannot open dataset " + path);
    Hid ft(h.dget_type(d.id), h.tclose);
    if (h.tget_class(ft.id) != H5T_STRING_) throw Err(IFE_ERR_INVALID, "mesh file: " + path + " is not a string dataset");
    Hid sp(h.dget_space(d.id), h.sclose);
    const int nd = h.sget_ndims(sp.id);
    std::vector<hsize_t> dims(nd > 0 ? nd : 0);
    if (nd > 0) h.sget_dims(sp.id, dims.data(), nullptr);
    size_t n = 1;
    for (auto x : dims) n *= (size_t)x;
    std::vector<std::string> out;
    Hid mt(h.tcopy(*h.t_c_s1), h.tclose);
    h.tset_cset(mt.id, h.tget_cset(ft.id));  // HDF5 does not convert between ASCII and UTF-8
    if (h.tis_variable_str(ft.id) > 0) {
        h.tset_size(mt.id, H5T_VARIABLE_);
        std::vector<char*> buf(n, nullptr);
        if (n && h.dread(d.id, mt.id, H5S_ALL_, H5S_ALL_, H5P_DEFAULT_, buf.data()) < 0)
            throw Err(IFE_ERR_INVALID, "mesh file: cannot read " + path);
        for (auto* c : buf) out.emplace_back(c ? c : "");
        if (n) h.dvlen_reclaim(mt.id, sp.id, H5P_DEFAULT_, buf.data());
    } else {
        const size_t sz = h.tget_size(ft.id);
        h.tset_size(mt.id, sz);
        std::vector<char> buf(n * sz + 1, 0);
        if (n && h.dread(d.id, mt.id, H5S_ALL_, H5S_ALL_, H5P_DEFAULT_, buf.data()) < 0)
            throw Err(IFE_ERR_INVALID, "mesh file: cannot read " + path);
        for (size_t i = 0; i < n; i++) out.emplace_back(std::string(&buf[i * sz], strnlen(&buf[i * sz], sz)));
    }
    return out;
}

void write_num(hid_t f, const std::string& path, hid_t memtype, const std::vector<hsize_t>& dims, const void* data) {
    const H5& h = h5();
    Hid sp(h.screate_simple((int)dims.size(), dims.data(), nullptr), h.sclose);
    Hid d(h.dcreate2(f, path.c_str(), memtype, sp.id, H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_), h.dclose);
    if (d.id < 0 || h.dwrite(d.id, memtype, H5S_ALL_, H5S_ALL_, H5P_DEFAULT_, data) < 0)
        throw Err(IFE_ERR_INVALID, "mesh file: cannot write " + path);
}

void write_strings(hid_t f, const std::string& path, const std::vector<std::string>& s) {
    // MATLAB's 'string' datatype: variable-length UTF-8 strings
    const H5& h = h5();
    Hid mt(h.tcopy(*h.t_c_s1), h.tclose);
    h.tset_size(mt.id, H5T_VARIABLE_);
    h.tset_cset(mt.id, H5T_CSET_UTF8_);
    std::vector<const char*> p;
    for (auto& x : s) p.push_back(x.c_str());
    write_num(f, path, mt.id, {(hsize_t)s.size()}, p.data());
}

void make_group(hid_t f, const std::string& path) {
    const H5& h = h5();
    if (exists(f, path)) return;
    Hid g(h.gcreate2(f, path.c_str(), H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_), h.gclose);
    if (g.id < 0) throw Err(IFE_ERR_INVALID, "mesh file: cannot create group " + path);
}

template <class F>
int guard(F&& f) {
    try {
        f();
        return IFE_OK;
    } catch (const Err& e) {
        ife::set_last_error(e.what());
        return e.code;
    } catch (const std::invalid_argument& e) {  // no HDF5 library
        ife::set_last_error(e.what());
        return IFE_ERR_UNSUPPORTED;
    } catch (const std::bad_alloc&) {
        ife::set_last_error("host allocation failed");
        return IFE_ERR_NOMEM;
    } catch (const std::exception& e) {
        ife::set_last_error(e.what());
        return IFE_ERR_INVALID;
    }
}

int dim_of_type(const std::string& t) {
    if (t == "B2_2") return 2;
    if (t == "NURBS3_Cube2") return 3;
    return 0;
}

}  // namespace

struct ife_mesh_file {
    int32_t dim = 0, ipcount1D = 3;
    int64_t n_nodes = 0;
    std::vector<double> coords;   // [node][dim]
    std::vector<double> weights;  // [node] ("W" node data, ones without it)
    bool has_weights = false;
    struct Group {
        std::string name, type;
        int64_t n_elems = 0;
        int32_t npe = 0;
        std::vector<int64_t> elems;  // [e][npe]
        std::vector<double> ext;     // [e][npe][npe], C[e][j][k] = ElemData(j,k)
    };
    std::vector<Group> egroups;
    std::vector<std::string> ngroup_names;
    std::vector<std::vector<int64_t>> ngroups;
};

extern "C" {

int ife_mesh_load_hdf5(const char* path, int32_t ipcount1D, ife_mesh_file** out) {
    return guard([&] {
        if (!path || !out) throw Err(IFE_ERR_INVALID, "null argument");
        *out = nullptr;
        const H5& h = h5();
        Hid f(h.fopen(path, H5F_ACC_RDONLY_, H5P_DEFAULT_), h.fclose);
        if (f.id < 0) throw Err(IFE_ERR_INVALID, std::string("cannot open mesh file ") + path);
        auto m = new ife_mesh_file();
        std::unique_ptr<ife_mesh_file> keep(m);
        m->ipcount1D = ipcount1D;
        // Mesh::LoadNodesFromFile, single rank (mesh.cpp:91-126)
        std::vector<hsize_t> dims;
        std::vector<double> nodes = read_num<double>(f.id, "/nodes", *h.t_double, dims);
        if (dims.size() != 2 || (dims[0] != 2 && dims[0] != 3))
            throw Err(IFE_ERR_INVALID, "/nodes must be [dim][n_nodes] with dim 2 or 3");
        m->dim = (int32_t)dims[0];
        m->n_nodes = (int64_t)dims[1];
        m->coords.resize((size_t)m->n_nodes * m->dim);
        for (int64_t i = 0; i < m->n_nodes; i++)
            for (int d = 0; d < m->dim; d++) m->coords[i * m->dim + d] = nodes[(size_t)d * m->n_nodes + i];
        m->weights.assign((size_t)m->n_nodes, 1.0);
        if (exists(f.id, "/node_DataTypes")) {
            auto names = read_strings(f.id, "/node_DataTypes");
            std::vector<double> nd = read_num<double>(f.id, "/node_Data", *h.t_double, dims);
            const size_t nn = dims.empty() ? 0 : (size_t)dims.back();
            if (nn != (size_t)m->n_nodes || nd.size() != names.size() * nn)
                throw Err(IFE_ERR_INVALID, "/node_Data must be [n_data][n_nodes]");
            for (size_t k = 0; k < names.size(); k++)
                if (names[k] == "W") {
                    for (size_t i = 0; i < nn; i++) m->weights[i] = nd[k * nn + i];
                    m->has_weights = true;
                }
        }
        // Mesh::LoadNodeGroupsFromFile (mesh.cpp:193-208)
        if (exists(f.id, "/nodegroupnames")) {
            m->ngroup_names = read_strings(f.id, "/nodegroupnames");
            for (auto& nm : m->ngroup_names) {
                std::vector<uint64_t> g = read_num<uint64_t>(f.id, "/nodegroups/" + nm, *h.t_uint64, dims);
                std::vector<int64_t> v(g.begin(), g.end());
                for (auto x : v)
                    if (x < 0 || x >= m->n_nodes) throw Err(IFE_ERR_INVALID, "node group " + nm + ": node id out of range");
                m->ngroups.push_back(std::move(v));
            }
        }
        // Mesh::LoadElementGroupsFromFile (mesh.cpp:237-262) + elementgroup::AddElems
        auto names = read_strings(f.id, "/elementgroupnames");
        auto types = read_strings(f.id, "/elementgrouptypes");
        if (names.size() != types.size()) throw Err(IFE_ERR_INVALID, "element group names and types differ in count");
        for (size_t g = 0; g < names.size(); g++) {
            ife_mesh_file::Group G;
            G.name = names[g];
            G.type = types[g];
            std::vector<uint64_t> el = read_num<uint64_t>(f.id, "/elementgroups/" + G.name, *h.t_uint64, dims);
            if (dims.size() != 2) throw Err(IFE_ERR_INVALID, "/elementgroups/" + G.name + " must be [nodes][elements]");
            G.npe = (int32_t)dims[0];
            G.n_elems = (int64_t)dims[1];
            G.elems.resize((size_t)G.n_elems * G.npe);
            for (int64_t e = 0; e < G.n_elems; e++)
                for (int j = 0; j < G.npe; j++) {
                    const uint64_t v = el[(size_t)j * G.n_elems + e];
                    if (v >= (uint64_t)m->n_nodes) throw Err(IFE_ERR_INVALID, "element group " + G.name + ": node id out of range");
                    G.elems[(size_t)e * G.npe + j] = (int64_t)v;
                }
            const std::string dp = "/elementgroups/" + G.name + "_Data";
            if (exists(f.id, dp)) {
                std::vector<double> ed = read_num<double>(f.id, dp, *h.t_double, dims);
                const int n = G.npe;
                if (dims.size() != 3 || dims[0] != (hsize_t)n || dims[1] != (hsize_t)n || dims[2] != (hsize_t)G.n_elems)
                    throw Err(IFE_ERR_INVALID, dp + " must be [n][n][elements]");
                G.ext.resize((size_t)G.n_elems * n * n);
                for (int64_t e = 0; e < G.n_elems; e++)
                    for (int j = 0; j < n; j++)
                        for (int k = 0; k < n; k++)
                            G.ext[((size_t)e * n + j) * n + k] = ed[((size_t)k * n + j) * G.n_elems + e];
            }
            m->egroups.push_back(std::move(G));
        }
        *out = keep.release();
    });
}

void ife_mesh_file_free(ife_mesh_file* m) { delete m; }

int ife_mesh_file_info(const ife_mesh_file* m, int32_t* dim, int64_t* n_nodes, int32_t* n_element_groups,
                       int32_t* n_node_groups) {
    return guard([&] {
        if (!m) throw Err(IFE_ERR_INVALID, "null mesh file");
        if (dim) *dim = m->dim;
        if (n_nodes) *n_nodes = m->n_nodes;
        if (n_element_groups) *n_element_groups = (int32_t)m->egroups.size();
        if (n_node_groups) *n_node_groups = (int32_t)m->ngroups.size();
    });
}

int ife_mesh_file_element_group(const ife_mesh_file* m, int32_t group, const char** name, const char** type,
                                int64_t* n_elems, int32_t* nodes_per_elem) {
    return guard([&] {
        if (!m) throw Err(IFE_ERR_INVALID, "null mesh file");
        if (group < 0 || group >= (int32_t)m->egroups.size()) throw Err(IFE_ERR_INVALID, "element group out of range");
        const auto& G = m->egroups[group];
        if (name) *name = G.name.c_str();
        if (type) *type = G.type.c_str();
        if (n_elems) *n_elems = G.n_elems;
        if (nodes_per_elem) *nodes_per_elem = G.npe;
    });
}

int ife_mesh_file_node_group(const ife_mesh_file* m, int32_t group, const char** name, int64_t* n_nodes,
                             const int64_t** nodes) {
    return guard([&] {
        if (!m) throw Err(IFE_ERR_INVALID, "null mesh file");
        if (group < 0 || group >= (int32_t)m->ngroups.size()) throw Err(IFE_ERR_INVALID, "node group out of range");
        if (name) *name = m->ngroup_names[group].c_str();
        if (n_nodes) *n_nodes = (int64_t)m->ngroups[group].size();
        if (nodes) *nodes = m->ngroups[group].data();
    });
}

int ife_mesh_file_desc(const ife_mesh_file* m, int32_t group, ife_mesh_desc* out) {
    return guard([&] {
        if (!m || !out) throw Err(IFE_ERR_INVALID, "null argument");
        if (group < 0 || group >= (int32_t)m->egroups.size()) throw Err(IFE_ERR_INVALID, "element group out of range");
        const auto& G = m->egroups[group];
        const int d = dim_of_type(G.type);
        if (d == 0) throw Err(IFE_ERR_UNSUPPORTED, "element type " + G.type + " (the engine runs B2_2 and NURBS3_Cube2)");
        if (d != m->dim) throw Err(IFE_ERR_INVALID, "element type " + G.type + " does not match the node dimension");
        if (G.ext.empty()) throw Err(IFE_ERR_INVALID, "element group " + G.name + " has no extraction operators (_Data)");
        std::memset(out, 0, sizeof(*out));
        out->dim = m->dim;
        out->ipcount1D = m->ipcount1D;
        out->n_nodes = m->n_nodes;
        out->coords = m->coords.data();
        out->weights = m->weights.data();
        out->n_elems = G.n_elems;
        out->nodes_per_elem = G.npe;
        out->elems = G.elems.data();
        out->extraction = G.ext.data();
        out->extraction_1d = nullptr;
    });
}

int ife_mesh_save_hdf5(const char* path, const ife_mesh_desc* mesh, const char* group_name, int32_t n_node_groups,
                       const char* const* node_group_names, const int64_t* node_group_ptr, const int64_t* node_group_nodes) {
    return guard([&] {
        if (!path || !mesh || !group_name) throw Err(IFE_ERR_INVALID, "null argument");
        if (mesh->dim != 2 && mesh->dim != 3) throw Err(IFE_ERR_INVALID, "dim must be 2 or 3");
        const int n = mesh->nodes_per_elem;
        if (n != (mesh->dim == 2 ? 9 : 27)) throw Err(IFE_ERR_INVALID, "nodes_per_elem must be 9 (2D) or 27 (3D)");
        if (!mesh->coords || !mesh->elems || (!mesh->extraction && !mesh->extraction_1d))
            throw Err(IFE_ERR_INVALID, "mesh arrays missing");
        if (n_node_groups > 0 && (!node_group_names || !node_group_ptr || !node_group_nodes))
            throw Err(IFE_ERR_INVALID, "node group arrays missing");
        const H5& h = h5();
        Hid f(h.fcreate(path, H5F_ACC_TRUNC_, H5P_DEFAULT_, H5P_DEFAULT_), h.fclose);
        if (f.id < 0) throw Err(IFE_ERR_INVALID, std::string("cannot create mesh file ") + path);
        const int64_t NN = mesh->n_nodes, NE = mesh->n_elems;
        const int dim = mesh->dim;
        std::vector<double> nodes((size_t)dim * NN);
        for (int64_t i = 0; i < NN; i++)
            for (int d = 0; d < dim; d++) nodes[(size_t)d * NN + i] = mesh->coords[i * dim + d];
        write_num(f.id, "/nodes", *h.t_double, {(hsize_t)dim, (hsize_t)NN}, nodes.data());
        write_strings(f.id, "/node_DataTypes", {"W"});
        std::vector<double> w((size_t)NN, 1.0);
        if (mesh->weights) w.assign(mesh->weights, mesh->weights + NN);
        write_num(f.id, "/node_Data", *h.t_double, {1, (hsize_t)NN}, w.data());
        std::vector<std::string> ngn;
        for (int g = 0; g < n_node_groups; g++) ngn.emplace_back(node_group_names[g]);
        write_strings(f.id, "/nodegroupnames", ngn);
        make_group(f.id, "/nodegroups");
        for (int g = 0; g < n_node_groups; g++) {
            std::vector<uint64_t> v(node_group_nodes + node_group_ptr[g], node_group_nodes + node_group_ptr[g + 1]);
            write_num(f.id, "/nodegroups/" + ngn[g], *h.t_uint64, {(hsize_t)v.size()}, v.data());
        }
        write_strings(f.id, "/elementgroupnames", {group_name});
        write_strings(f.id, "/elementgrouptypes", {dim == 2 ? "B2_2" : "NURBS3_Cube2"});
        make_group(f.id, "/elementgroups");
        std::vector<uint64_t> el((size_t)n * NE);
        for (int64_t e = 0; e < NE; e++)
            for (int j = 0; j < n; j++) el[(size_t)j * NE + e] = (uint64_t)mesh->elems[e * n + j];
        const std::string gp = std::string("/elementgroups/") + group_name;
        write_num(f.id, gp, *h.t_uint64, {(hsize_t)n, (hsize_t)NE}, el.data());
        std::vector<double> ed((size_t)n * n * NE);
        for (int64_t e = 0; e < NE; e++) {
            double C[27 * 27];
            if (mesh->extraction) {
                std::memcpy(C, mesh->extraction + (size_t)e * n * n, sizeof(double) * n * n);
            } else {  // C = kron(Cz, kron(Cy, Cx)) from the 1D operators (bezierExtraction3D.m:17-27)
                const double* c1 = mesh->extraction_1d + (size_t)e * dim * 9;
                for (int j = 0; j < n; j++)
                    for (int k = 0; k < n; k++) {
                        double v = 1.0;
                        int jj = j, kk = k;
                        for (int d = 0; d < dim; d++) {
                            v *= c1[d * 9 + (jj % 3) * 3 + (kk % 3)];
                            jj /= 3;
                            kk /= 3;
                        }
                        C[j * n + k] = v;
                    }
            }
            for (int j = 0; j < n; j++)
                for (int k = 0; k < n; k++) ed[((size_t)k * n + j) * NE + e] = C[j * n + k];
        }
        write_num(f.id, gp + "_Data", *h.t_double, {(hsize_t)n, (hsize_t)n, (hsize_t)NE}, ed.data());
        // one-rank partition (SaveToHDF.m:54-83): everything on rank 0, no ghosts
        uint64_t ncores = 1;
        write_num(f.id, "/NCores", *h.t_uint64, {1}, &ncores);
        make_group(f.id, "/Partition");
        make_group(f.id, "/Partition/0");
        uint64_t nr[2] = {0, (uint64_t)(NN - 1)};
        write_num(f.id, "/Partition/0/noderange", *h.t_uint64, {2}, nr);
        uint8_t no = 0, yes = 1;
        write_num(f.id, "/Partition/0/Hasghosts", *h.t_uint8, {1}, &no);
        write_num(f.id, "/Partition/0/Haselems", *h.t_uint8, {1}, &yes);
        uint64_t er[2] = {0, (uint64_t)(NE - 1)};
        write_num(f.id, "/Partition/0/elemrange", *h.t_uint64, {2, 1}, er);
        std::vector<uint8_t> hng((size_t)n_node_groups, 1);
        std::vector<uint64_t> ngr((size_t)2 * n_node_groups);
        for (int g = 0; g < n_node_groups; g++) {
            ngr[g] = 0;
            ngr[n_node_groups + g] = (uint64_t)(node_group_ptr[g + 1] - node_group_ptr[g]) - 1;
        }
        write_num(f.id, "/Partition/0/hasnodegroup", *h.t_uint8, {(hsize_t)n_node_groups}, hng.data());
        write_num(f.id, "/Partition/0/nodegrouprange", *h.t_uint64, {2, (hsize_t)n_node_groups}, ngr.data());
    });
}

}  // extern "C"
