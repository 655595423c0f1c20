// Zero-mode checkpoint files (host side): the reference's V2/V3/V4 format,
// byte for byte, and the exact-resume companion file of this library.
//
// Format (pf_zero_mode_checkpoint.cpp:20-118, 284-483, 485-760): a fixed
// little-endian header (V2 512 B, V3 800 B, V4 904 B) holding the grid, the
// time step, the zero-mode runtime scalars, the run's provenance selectors
// (64-byte NUL-padded strings) and identities (96-byte), two fingerprints and
// the FNV-1a 64 checksum of (header with a zero checksum field, phi, Y, x_B,
// dY/dt as FP64, then in V4 the elastic warm state u(k) as complex FP32 over
// the full half spectrum [3][Nx][Ny][Nz/2+1]).  Written to <path>.tmp and
// renamed when complete; read back only when the caller's expected
// provenance matches and the checksum holds.
//
// Exact-resume companion (<path>.state, no reference counterpart): the FP64
// warm state in the library's compact spectral layout and the warm-start
// history, bound to the V4 file by its payload checksum.  With it a restart
// continues bit for bit; without it the restart starts from the V4 FP32 warm
// state, as the reference does.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace pf {
namespace ckpt {

constexpr int kSelectorBytes = 64;
constexpr int kIdentityBytes = 96;
constexpr std::uint32_t kHeaderV2 = 512, kHeaderV3 = 800, kHeaderV4 = 904;

// Selector and identity values the reference run writes (pf_zero_mode_checkpoint.h:11-22)
extern const char* const kModeV1;
extern const char* const kHostBackendV1;
extern const char* const kExplicitContextNV1;
extern const char* const kReactionTangentNV1;
extern const char* const kLegacyUnspecified;
extern const char* const kElasticNotRequired;
extern const char* const kElasticWarmStartResidualV1;

struct Provenance {
  std::string zero_mode, backend, composition_mode, y_update_mode, explicit_context,
      reaction_discretization;
  std::string initial_state_class, fixture_manifest_sha256, profile_library_manifest_sha256;
  std::string elastic_solver_mode;
  std::uint64_t elastic_solver_fingerprint = 0, parameter_fingerprint = 0;
};

struct Checkpoint {
  std::uint32_t version = 0;   // as read (2, 3, 4); the writer picks 3 or 4
  std::uint64_t accepted_step = 0;
  int nx = 0, ny = 0, nz = 0;
  double dt_code = 0.0, temperature_K = 0.0;
  Provenance prov;
  // zero-mode runtime
  double target_mass_code = 0.0, last_lambda = 0.0, last_residual_code = 0.0, last_derivative_code = 0.0;
  std::uint64_t last_iterations = 0, accepted_zero_mode_steps = 0;
  // elastic warm state (V4)
  bool elastic_present = false;
  std::uint64_t elastic_source_field_step = 0, elastic_last_iterations = 0;
  double elastic_last_relative_residual = 0.0;
  std::vector<float> displacement_k;   // 6 * nx*ny*(nz/2+1): ux, uy, uz as (re, im) pairs
  std::vector<double> phi, Y, xB, dY_dt_prev;
  std::uint64_t payload_checksum = 0;  // as written / read
};

std::uint64_t fnv1a64(const void* data, std::size_t size, std::uint64_t seed = 1469598103934665603ULL);
bool elastic_expected(const Provenance& p);
bool validate(const Checkpoint& c, std::string* err);
bool write(const std::string& path, Checkpoint& c, std::string* err);   // sets c.payload_checksum
bool read(const std::string& path, const Provenance& expected, Checkpoint* out, std::string* err);

// exact-resume companion
struct ExactState {
  std::uint64_t bound_checksum = 0;   // payload checksum of the V4 file it belongs to
  std::uint64_t accepted_step = 0;
  std::int32_t precision = 0;         // the handle's precision
  std::int32_t nstates = 0;           // 1 + history length
  std::uint64_t n = 0;                // complex values per component (compact layout)
  std::vector<double> states;         // nstates * 3 * n complex (re, im): current, then history newest first
  // mixed precision: the unclamped phi* of the phi update ([Nx][Ny][Nz]) and
  // its retained spectrum (n complex), or empty
  std::vector<double> phi_star, phi_spec;
  // the handle's tile plan (z-pass budget, y and x column-tile indices), so a
  // restart in another process transforms with the same tiles; -1: absent
  // (files of format versions 1 and 2)
  std::int32_t plan[3] = {-1, -1, -1};
};
bool write_exact(const std::string& path, const ExactState& s, std::string* err);
// false with an empty error when the file does not exist
bool read_exact(const std::string& path, ExactState* s, std::string* err);

}  // namespace ckpt
}  // namespace pf
