// Drives the C++ host mirror (curvami-official_amd/host/curvami_gpu.hpp) the way
// the reference's main drives Meshless: nodes, laminate, creases, domains,
// folds, recomputeShapeFunctions, initAssembly, assembleSystem. Input and
// output are raw little-endian arrays in a directory (written and checked by
// tests/test_gpu_host_mirror.py).
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../curvami-official_amd/host/curvami_gpu.hpp"

template <typename T>
static std::vector<T> load(const std::string& path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("cannot open " + path);
    const std::streamsize n = f.tellg();
    f.seekg(0);
    std::vector<T> v(static_cast<size_t>(n) / sizeof(T));
    f.read(reinterpret_cast<char*>(v.data()), n);
    return v;
}

template <typename T>
static void save(const std::string& path, const std::vector<T>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
}

int main(int argc, char** argv) {
    if (argc < 2) { std::cerr << "usage: mirror_check <dir>\n"; return 2; }
    const std::string d = argv[1];
    try {
        const auto meta = load<double>(d + "/meta.bin");       // rho, order, Kt, Ang0, flagRig
        curvami_gpu::Meshless plate(static_cast<int>(meta[1]));
        plate.rho = meta[0];
        plate.setNodes(load<double>(d + "/X.bin"), load<double>(d + "/Y.bin"));
        plate.setLaminate(load<double>(d + "/plies.bin"));
        const auto cr = load<double>(d + "/crease.bin");      // nv x, nv y, ext_s, ext_e
        const size_t nv = (cr.size() - 2) / 2;
        plate.addCrease(std::vector<double>(cr.begin(), cr.begin() + nv),
                        std::vector<double>(cr.begin() + nv, cr.begin() + 2 * nv), cr[2 * nv] != 0, cr[2 * nv + 1] != 0);
        curvami_gpu::GaussPoints face;
        face.x = load<double>(d + "/fx.bin"); face.y = load<double>(d + "/fy.bin"); face.w = load<double>(d + "/fw.bin");
        plate.addDomain(CVM_DOMAIN_FACE, face);
        auto gp = [&](const char* p) {
            curvami_gpu::GaussPoints g;
            const std::string b = d + "/" + p;
            g.x = load<double>(b + "x.bin"); g.y = load<double>(b + "y.bin"); g.w = load<double>(b + "w.bin");
            g.nx = load<double>(b + "nx.bin"); g.ny = load<double>(b + "ny.bin");
            g.tx = load<double>(b + "tx.bin"); g.ty = load<double>(b + "ty.bin");
            return g;
        };
        plate.addFold(gp("L"), gp("R"), meta[2], meta[3]);
        // boundary domains (point / edge) with their BoundaryCondition
        const int nb = meta.size() > 5 ? static_cast<int>(meta[5]) : 0;
        for (int k = 0; k < nb; ++k) {
            const std::string b = "b" + std::to_string(k);
            const auto bm = load<double>(d + "/" + b + "meta.bin");   // kind, active, val[4], force[3], moment
            const int dom = plate.addDomain(static_cast<int>(bm[0]), gp(b.c_str()));
            plate.setBoundaryCondition(dom, static_cast<int>(bm[1]), &bm[2], &bm[6], bm[9]);
        }
        plate.recomputeShapeFunctions();
        plate.recomputeShapeFunctions();   // again: the device search and every domain in one call
        plate.initAssembly();
        curvami_gpu::CSC K;
        std::vector<double> F;
        plate.assembleSystem(K, F, load<double>(d + "/x.bin"), meta[4] != 0);
        save(d + "/K_outer.bin", K.outer);
        save(d + "/K_inner.bin", K.inner);
        save(d + "/K_values.bin", K.values);
        save(d + "/F.bin", F);
        save(d + "/map0.bin", plate.csrMapping(0));
        if (nb > 0) {
            std::vector<double> c;
            std::vector<curvami_gpu::Meshless::Triplet> t;
            int nrows = 0;
            plate.evaluateConstraints(load<double>(d + "/x.bin"), c, t, nrows, meta[4] != 0);
            std::vector<int> rows, cols;
            std::vector<double> vals;
            for (const auto& e : t) { rows.push_back(e.row); cols.push_back(e.col); vals.push_back(e.value); }
            save(d + "/c.bin", c);
            save(d + "/J_rows.bin", rows);
            save(d + "/J_cols.bin", cols);
            save(d + "/J_vals.bin", vals);
            save(d + "/F_ext.bin", plate.computeExternalForce());
        }
        std::printf("mirror_check: nnz=%zu\n", K.values.size());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "mirror_check error: %s\n", e.what());
        return 1;
    }
    return 0;
}
