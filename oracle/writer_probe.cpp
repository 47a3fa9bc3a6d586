// Test infrastructure only: writes the VTK and CSV results of a parsed .fem model
// filled with synthetic values, so that the reference's writer (built with
// -DPROBE_REFERENCE against the reference sources, src/vtk_writer.cpp) and this
// repository's writer (lightfem_amd/csrc/host/vtk_writer.cpp) can be compared byte for
// byte on identical inputs.
//
//   writer_probe <model.fem> <output prefix> <seed>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <exception>
#include <string>

#ifdef PROBE_REFERENCE
#include "lightfem/input_parser.hpp"
#include "lightfem/solver.hpp"
#include "lightfem/vtk_writer.hpp"
#else
#include "lightfem/lightfem.hpp"
#endif

namespace {

std::uint64_t state = 0;

std::uint64_t next() {  // splitmix64
    std::uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Doubles over a wide range of magnitudes, both signs, with exact zeros, negative
// zero, integers, tiny and huge values and subnormals among them.
double value() {
    const std::uint64_t r = next();
    const double u = static_cast<double>(r >> 11) * 0x1.0p-53;
    switch (r & 15) {
        case 0: return 0.0;
        case 1: return -0.0;
        case 2: return std::floor((u - 0.5) * 2000.0);
        case 3: return (u - 0.5) * 1e-310;
        case 4: return (u - 0.5) * 1e300;
        default: return (u - 0.5) * std::pow(10.0, static_cast<double>(static_cast<int>((r >> 4) % 41) - 20));
    }
}

// FNV-1a over the parsed model's records (ids, coordinates, connectivity, material
// values, supports, loads, source lines), so two parsers can be compared exactly.
struct Digest {
    std::uint64_t h = 1469598103934665603ULL;
    template <typename T>
    void add(const T& v) {
        const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
        for (std::size_t i = 0; i < sizeof(T); ++i) h = (h ^ p[i]) * 1099511628211ULL;
    }
};

std::uint64_t model_digest(const lightfem::Model& m) {
    Digest d;
    for (const auto& n : m.nodes) {
        d.add(n.id); d.add(n.position.x); d.add(n.position.y); d.add(n.source_line);
    }
    for (const auto& x : m.materials) {
        d.add(x.id); d.add(x.young_modulus); d.add(x.poisson_ratio); d.add(x.thickness); d.add(x.source_line);
    }
    for (const auto& e : m.elements) {
        d.add(e.id); d.add(e.node_ids); d.add(e.material_id); d.add(e.source_line);
    }
    for (const auto& t : m.trusses) {
        d.add(t.id); d.add(t.node_ids); d.add(t.material_id); d.add(t.area); d.add(t.source_line);
    }
    for (const auto& b : m.boundaries) {
        d.add(b.node_id); d.add(b.ux.has_value()); d.add(b.ux.value_or(0.0)); d.add(b.uy.has_value());
        d.add(b.uy.value_or(0.0)); d.add(b.source_line);
    }
    for (const auto& l : m.loads) {
        d.add(l.node_id); d.add(l.fx); d.add(l.fy); d.add(l.source_line);
    }
    d.add(m.settings.tolerance); d.add(m.settings.max_iterations); d.add(m.settings.deformation_scale);
    return d.h;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: writer_probe <model.fem> <output prefix> <seed>\n");
        return 1;
    }
    try {
        state = std::stoull(argv[3]);
        lightfem::Model model = lightfem::InputParser{}.parse_file(argv[1]);
        std::printf("model_digest %016llx nodes %zu elements %zu materials %zu\n",
                    static_cast<unsigned long long>(model_digest(model)), model.nodes.size(), model.elements.size(),
                    model.materials.size());
        model.sort_and_index();
        lightfem::AnalysisResult result;
        const std::size_t dofs = model.summary().dofs;
        result.displacement.resize(dofs);
        result.reaction.resize(dofs);
        for (double& v : result.displacement) v = value();
        for (double& v : result.reaction) v = value();
        for (const lightfem::Node& node : model.nodes) {
            lightfem::NodeResult nr;
            nr.node_id = node.id;
            for (double& v : nr.stress) v = value();
            nr.von_mises = value();
            result.node_results.push_back(nr);
        }
        auto field = [](lightfem::ElementField& f) {
            for (double& v : f.strain) v = value();
            for (double& v : f.stress) v = value();
            f.von_mises = value();
        };
        for (const lightfem::Q4Element& e : model.elements) {
            lightfem::ElementResult er;
            er.element_id = e.id;
            field(er.field);
            for (auto& gp : er.gauss_points) {
                gp.xi = value();
                gp.eta = value();
                field(gp.field);
            }
            result.element_results.push_back(er);
        }
        for (const lightfem::TrussElement& t : model.trusses) {
            lightfem::ElementResult er;
            er.element_id = t.id;
            field(er.field);
            result.element_results.push_back(er);
        }
        const std::string prefix = argv[2];
        lightfem::write_vtk(model, result, prefix + ".vtk");
        (void)lightfem::write_csv_results(model, result, prefix);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
    return 0;
}
