// Test infrastructure: a command-line driver around the reference's own
// checkpoint writer and reader (pf_zero_mode_checkpoint.cpp, compiled from
// its location in the reference tree by oracle/Makefile into oracle/_ref/).
// The checkpoint tests use it to check this library's files against the
// reference implementation in both directions.  Nothing in the product
// links or runs it.
//
//   ckpt_ref_tool roundtrip IN OUT PROV...   read IN with the reference
//       reader (expected provenance PROV), write it back with the reference
//       writer to OUT
//   ckpt_ref_tool make OUT NX NY NZ ELASTIC PROV...   write a checkpoint of
//       deterministic content (the values make_value() defines) with the
//       reference writer
//
// PROV is 12 arguments: zero_mode backend composition_mode y_update_mode
// explicit_context reaction_discretization elastic_solver_mode
// initial_state_class fixture_manifest_sha256 profile_library_manifest_sha256
// parameter_fingerprint(hex) elastic_solver_fingerprint(hex).
#include <cstdio>
#include <cstdlib>
#include <string>

#include "pf_zero_mode_checkpoint.h"

namespace {

pf_zero_mode::Provenance parse_prov(char** a) {
  pf_zero_mode::Provenance p;
  p.zero_mode = a[0];
  p.backend = a[1];
  p.composition_mode = a[2];
  p.y_update_mode = a[3];
  p.explicit_context = a[4];
  p.reaction_discretization = a[5];
  p.elastic_solver_mode = a[6];
  p.initial_state_class = a[7];
  p.fixture_manifest_sha256 = a[8];
  p.profile_library_manifest_sha256 = a[9];
  p.parameter_fingerprint = std::strtoull(a[10], nullptr, 16);
  p.elastic_solver_fingerprint = std::strtoull(a[11], nullptr, 16);
  return p;
}

// deterministic content: array a (0 phi, 1 Y, 2 xB, 3 dYdt, 4 displacement), index i
double make_value(int a, std::size_t i) {
  const double t = static_cast<double>(i);
  switch (a) {
    case 0: return 0.5 + 0.5 * ((i * 37u) % 101u) / 101.0 - 0.25;
    case 1: return -4.0 + 0.01 * ((i * 13u) % 97u);
    case 2: return 0.02 + 0.0001 * ((i * 7u) % 89u);
    case 3: return 1.0e-6 * (t - 10.0);
    default: return 0.001 * ((i * 11u) % 83u) - 0.04;
  }
}

}  // namespace

int main(int argc, char** argv) {
  std::string err;
  if (argc == 4 + 12 && std::string(argv[1]) == "roundtrip") {
    pf_zero_mode::Checkpoint c;
    if (!pf_zero_mode::read_checkpoint(argv[2], parse_prov(argv + 4), &c, &err)) {
      std::printf("READ_FAIL %s\n", err.c_str());
      return 1;
    }
    if (!pf_zero_mode::write_checkpoint(argv[3], c, &err)) {
      std::printf("WRITE_FAIL %s\n", err.c_str());
      return 1;
    }
    std::printf("OK\n");
    return 0;
  }
  if (argc == 7 + 12 && std::string(argv[1]) == "make") {
    pf_zero_mode::Checkpoint c;
    c.nx = std::atoi(argv[3]);
    c.ny = std::atoi(argv[4]);
    c.nz = std::atoi(argv[5]);
    const bool elastic = std::atoi(argv[6]) != 0;
    c.provenance = parse_prov(argv + 7);
    c.accepted_step = 9;
    c.dt_code = 0.01;
    c.temperature_K = 653.15;
    c.zero_mode.target_mass_code = 123.25;
    c.zero_mode.last_lambda = -2.5e-7;
    c.zero_mode.last_residual_code = 1.5e-14;
    c.zero_mode.last_derivative_code = 7.75;
    c.zero_mode.last_iterations = 3;
    c.zero_mode.accepted_zero_mode_steps = 9;
    const std::size_t n = static_cast<std::size_t>(c.nx) * c.ny * c.nz;
    std::vector<double>* f[4] = {&c.phi, &c.Y, &c.xB, &c.dY_dt_prev};
    for (int a = 0; a < 4; ++a)
      for (std::size_t i = 0; i < n; ++i) f[a]->push_back(make_value(a, i));
    if (elastic) {
      c.elastic.present = true;
      c.elastic.source_field_step = 8;
      c.elastic.last_iterations = 5;
      c.elastic.last_relative_residual = 4.5e-9;
      const std::size_t k = static_cast<std::size_t>(c.nx) * c.ny * (c.nz / 2 + 1);
      for (std::size_t i = 0; i < 6 * k; ++i)
        c.elastic.displacement_k.push_back(static_cast<float>(make_value(4, i)));
    }
    if (!pf_zero_mode::write_checkpoint(argv[2], c, &err)) {
      std::printf("WRITE_FAIL %s\n", err.c_str());
      return 1;
    }
    std::printf("OK\n");
    return 0;
  }
  std::fprintf(stderr, "usage: ckpt_ref_tool roundtrip IN OUT PROV... | make OUT NX NY NZ ELASTIC PROV...\n");
  return 2;
}
