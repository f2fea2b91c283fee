// ife_driver.cpp -- drives the engine through the C++ host wrapper (include/ife_model.hpp)
// the way IceFEM's Physics drives its models: create models, Setup, Assemble(step) for
// every staggered step, read K/f back.  Input/output are raw little-endian files written
// and read by tests/test_cpp_host.py.
//
// usage: ife_driver <problem.bin> <out.bin>
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "ife_model.hpp"

template <class T>
static T rd(std::ifstream& in) {
    T v;
    in.read(reinterpret_cast<char*>(&v), sizeof(T));
    if (!in) throw std::runtime_error("truncated problem file");
    return v;
}
template <class T>
static std::vector<T> rdv(std::ifstream& in, size_t n) {
    std::vector<T> v(n);
    in.read(reinterpret_cast<char*>(v.data()), n * sizeof(T));
    if (!in) throw std::runtime_error("truncated problem file");
    return v;
}
static std::string rds(std::ifstream& in) {
    int32_t n = rd<int32_t>(in);
    std::string s(n, ' ');
    in.read(&s[0], n);
    return s;
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::cerr << "usage: ife_driver problem.bin out.bin\n";
        return 2;
    }
    try {
        std::ifstream in(argv[1], std::ios::binary);
        int32_t dim = rd<int32_t>(in);
        int64_t NN = rd<int64_t>(in), NE = rd<int64_t>(in);
        int n = dim == 2 ? 9 : 27;
        auto coords = rdv<double>(in, NN * dim);
        auto W = rdv<double>(in, NN);
        auto elems = rdv<int64_t>(in, NE * n);
        auto ext = rdv<double>(in, NE * dim * 9);
        int32_t nf = rd<int32_t>(in);
        std::vector<std::string> names;
        std::vector<int32_t> steps;
        for (int i = 0; i < nf; i++) {
            names.push_back(rds(in));
            steps.push_back(rd<int32_t>(in));
        }
        ife_mesh_desc mesh{dim, 3, NN, coords.data(), W.data(), NE, n, elems.data(), nullptr, ext.data()};
        ife::Physics physics(mesh, names, steps);
        int32_t nmodels = rd<int32_t>(in);
        std::vector<ife::BaseModel*> models;
        for (int m = 0; m < nmodels; m++) {
            std::string type = rds(in), name = rds(in);
            auto mat = rd<ife_material>(in);
            auto pf = rd<ife_phasefield_opts>(in);
            auto opts = rd<ife_model_opts>(in);
            bool is_pf = type == "Fracture/PhaseField";
            models.push_back(&physics.CreateModel(type, name, mat, is_pf ? &pf : nullptr, is_pf ? &opts : nullptr));
        }
        double t = rd<double>(in), dt = rd<double>(in), du = rd<double>(in), ddu = rd<double>(in);
        physics.Setup();
        physics.SetTime(t, dt, du, ddu);
        int32_t nstate = rd<int32_t>(in);
        for (int i = 0; i < nstate; i++) {
            int32_t step = rd<int32_t>(in), kind = rd<int32_t>(in);
            int64_t len = rd<int64_t>(in);
            auto v = rdv<double>(in, len);
            if (hipMemcpy(physics.State(step, (ife_state_kind)kind), v.data(), len * 8, hipMemcpyHostToDevice) != hipSuccess)
                throw std::runtime_error("hipMemcpy");
        }
        int32_t nhist = rd<int32_t>(in);
        for (int i = 0; i < nhist; i++) {
            int32_t model = rd<int32_t>(in);
            std::string nm = rds(in);
            int64_t len = rd<int64_t>(in);
            auto v = rdv<double>(in, len);
            int64_t nc = 0, npc = 0;
            double* p = models[model]->History(nm, &nc, &npc);
            if (len > nc * npc) throw std::runtime_error("history too long");
            if (hipMemcpy(p, v.data(), len * 8, hipMemcpyHostToDevice) != hipSuccess) throw std::runtime_error("hipMemcpy");
        }
        std::ofstream out(argv[2], std::ios::binary);
        const int32_t ns = physics.NumSteps();
        out.write(reinterpret_cast<const char*>(&ns), 4);
        for (int32_t s = 0; s < ns; s++) {
            physics.Assemble(s);
            auto st = physics.Status();
            double *K, *f;
            physics.Kf(s, &K, &f);
            ife::Pattern p = physics.GetPattern(s);
            std::vector<double> hK(p.nnz), hf(p.n_rows);
            if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("sync");
            if (p.nnz && hipMemcpy(hK.data(), K, p.nnz * 8, hipMemcpyDeviceToHost) != hipSuccess) throw std::runtime_error("D2H");
            if (p.n_rows && hipMemcpy(hf.data(), f, p.n_rows * 8, hipMemcpyDeviceToHost) != hipSuccess) throw std::runtime_error("D2H");
            out.write(reinterpret_cast<const char*>(&p.nnz), 8);
            out.write(reinterpret_cast<const char*>(hK.data()), p.nnz * 8);
            out.write(reinterpret_cast<const char*>(&p.n_rows), 8);
            out.write(reinterpret_cast<const char*>(hf.data()), p.n_rows * 8);
            std::printf("step %d: %lld elements, %lld return-map iterations\n", s, (long long)st.elements_assembled,
                        (long long)st.return_map_iterations);
        }
        // Dirichlet constraints through the wrapper: AddConstraint's duplicate rule, then
        // Physics::Constraint and the residual norms on the device
        try {
            physics.SetConstraints(0, {1, 2, 1}, {0.0, 0.0, 1.0});
            std::cerr << "expected a doubly-constrained error\n";
            return 1;
        } catch (const ife::Error& e) {
            std::printf("doubly constrained dof rejected: %s\n", e.what());
        }
        physics.SetConstraints(0, {1, 2, 1}, {0.0, 0.0, 0.0});
        physics.Constraint(0);
        {
            int64_t nfree = 0, nnzc = 0;
            const int64_t* crp = nullptr;
            const int32_t* ccol = nullptr;
            double *Kc = nullptr, *fc = nullptr;
            physics.Constrained(0, &nfree, &nnzc, &crp, &ccol, &Kc, &fc);
            const ife::Pattern p0 = physics.GetPattern(0);
            if (nfree != p0.n_rows - 2 || nnzc <= 0 || nnzc >= p0.nnz) throw std::runtime_error("constrained system size");
            double* dx = nullptr;
            if (hipMalloc((void**)&dx, nfree * sizeof(double)) != hipSuccess) throw std::runtime_error("hipMalloc");
            std::vector<double> ones(nfree, 1.0);
            if (hipMemcpy(dx, ones.data(), nfree * 8, hipMemcpyHostToDevice) != hipSuccess) throw std::runtime_error("H2D");
            const auto e = physics.ResidualNorms(0, dx);
            (void)hipFree(dx);
            if (e[0] != 1.0 || !(e[1] >= 0) || e[1] != e[2]) throw std::runtime_error("residual norms");
            std::printf("constraints ok: %lld free dofs, %lld entries, |f_con|_1/n %.6e\n", (long long)nfree,
                        (long long)nnzc, e[1]);
        }
        models[0]->Commit(ife::CommitTypes::TIMEDEP);
        models[0]->ResetStep();
        // unknown model names raise like CreateModel (ModelRegister.cpp:72-75)
        try {
            ife_material m{};
            physics.CreateModel("Fracture/Unknown", "x", m);
            std::cerr << "expected an exception\n";
            return 1;
        } catch (const ife::Error& e) {
            std::printf("unknown model rejected: %s\n", e.what());
        }
        std::printf("driver ok\n");
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "ife_driver: " << e.what() << "\n";
        return 1;
    }
}
