// Zero-mode checkpoint files: serializer, reader, validation, fingerprints
// and the file-level C ABI.  The layout is described in pf_checkpoint.hpp.
#include "pf_checkpoint.hpp"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>

#include "../../include/pfcore.h"
#include "pf_error.hpp"

namespace pf {
namespace ckpt {

const char* const kModeV1 = "PF_CONSERVED_Y_ZERO_MODE_V1";
const char* const kHostBackendV1 = "HOST_NEWTON_BISECTION_V1";
const char* const kExplicitContextNV1 = "SM_EXPLICIT_CONTEXT_N_V1";
const char* const kReactionTangentNV1 = "SM_TANGENT_N_V1";
const char* const kLegacyUnspecified = "LEGACY_UNSPECIFIED";
const char* const kElasticNotRequired = "ELASTIC_RUNTIME_STATE_NOT_REQUIRED";
const char* const kElasticWarmStartResidualV1 = "ELASTIC_WARM_START_RESIDUAL_V1";

static_assert(sizeof(double) == 8 && sizeof(float) == 4, "checkpoint layout assumes IEEE binary64/32");

namespace {

bool set_err(std::string* err, const std::string& m) {
  if (err) *err = m;
  return false;
}

// Little-endian byte image of a header, built field by field in disk order.
struct Image {
  std::vector<unsigned char> b;
  template <typename V>
  void put(V v) {
    unsigned char t[sizeof(V)];
    std::memcpy(t, &v, sizeof(V));
    b.insert(b.end(), t, t + sizeof(V));
  }
  void put_text(const std::string& s, int width) {
    const size_t at = b.size();
    b.resize(at + width, 0);
    std::memcpy(b.data() + at, s.data(), s.size());
  }
};

// Sequential reader over a header image.
struct Cursor {
  const unsigned char* p;
  template <typename V>
  V get() {
    V v;
    std::memcpy(&v, p, sizeof(V));
    p += sizeof(V);
    return v;
  }
  // the stored text must equal `expected` and be NUL-terminated inside the field
  bool text_equals(const std::string& expected, int width) {
    const char* s = reinterpret_cast<const char*>(p);
    p += width;
    return expected.size() < (size_t)width && std::strncmp(s, expected.c_str(), width) == 0 &&
           s[expected.size()] == '\0';
  }
};

bool text_fits(const std::string& v, int width, const char* label, std::string* err) {
  if (v.empty() || v.size() >= (size_t)width) return set_err(err, std::string(label) + " is empty or too long");
  return true;
}

bool all_finite(const std::vector<double>& v) {
  for (double x : v)
    if (!std::isfinite(x)) return false;
  return true;
}
bool all_finite(const std::vector<float>& v) {
  for (float x : v)
    if (!std::isfinite(x)) return false;
  return true;
}

std::uint64_t k_count(int nx, int ny, int nz) {
  return (std::uint64_t)nx * (std::uint64_t)ny * (std::uint64_t)(nz / 2 + 1);
}

bool legacy_identity(const Provenance& p) {
  return p.initial_state_class == kLegacyUnspecified && p.fixture_manifest_sha256 == kLegacyUnspecified &&
         p.profile_library_manifest_sha256 == kLegacyUnspecified;
}

bool write_all(FILE* f, const void* d, size_t n) { return n == 0 || std::fwrite(d, 1, n, f) == n; }
bool read_all(FILE* f, void* d, size_t n) { return n == 0 || std::fread(d, 1, n, f) == n; }

}  // namespace

std::uint64_t fnv1a64(const void* data, std::size_t size, std::uint64_t seed) {
  const unsigned char* q = static_cast<const unsigned char*>(data);
  std::uint64_t h = seed;
  for (std::size_t i = 0; i < size; ++i) {
    h ^= (std::uint64_t)q[i];
    h *= 1099511628211ULL;
  }
  return h;
}

bool elastic_expected(const Provenance& p) {
  return p.elastic_solver_mode != kElasticNotRequired || p.elastic_solver_fingerprint != 0;
}

// pf_zero_mode_checkpoint.cpp:208-282
bool validate(const Checkpoint& c, std::string* err) {
  if (c.nx <= 0 || c.ny <= 0 || c.nz <= 0) return set_err(err, "checkpoint grid dimensions are invalid");
  const std::uint64_t n = (std::uint64_t)c.nx * (std::uint64_t)c.ny * (std::uint64_t)c.nz;
  if (n > (std::uint64_t)std::numeric_limits<std::size_t>::max())
    return set_err(err, "checkpoint element count overflows size_t");
  if (c.phi.size() != n || c.Y.size() != n || c.xB.size() != n || c.dY_dt_prev.size() != n)
    return set_err(err, "checkpoint field sizes do not match the grid");
  if (!(c.dt_code > 0.0) || !std::isfinite(c.dt_code) || !(c.temperature_K > 0.0) || !std::isfinite(c.temperature_K))
    return set_err(err, "checkpoint time/temperature is invalid");
  const Provenance& p = c.prov;
  if (p.zero_mode != kModeV1 || p.backend.empty() || p.composition_mode.empty() || p.y_update_mode.empty() ||
      p.explicit_context.empty() || p.reaction_discretization.empty() || p.initial_state_class.empty() ||
      p.fixture_manifest_sha256.empty() || p.profile_library_manifest_sha256.empty() || p.parameter_fingerprint == 0)
    return set_err(err, "checkpoint zero-mode provenance is incomplete");
  if (elastic_expected(p) != c.elastic_present)
    return set_err(err, "checkpoint elastic state/provenance presence mismatch");
  if (c.elastic_present) {
    if (p.elastic_solver_mode != kElasticWarmStartResidualV1 || p.elastic_solver_fingerprint == 0 ||
        c.accepted_step == 0 || c.elastic_source_field_step + 1 != c.accepted_step ||
        c.displacement_k.size() != 6 * k_count(c.nx, c.ny, c.nz) ||
        !std::isfinite(c.elastic_last_relative_residual) || c.elastic_last_relative_residual < 0.0 ||
        !all_finite(c.displacement_k))
      return set_err(err, "checkpoint elastic runtime state is invalid");
  } else if (!c.displacement_k.empty()) {
    return set_err(err, "checkpoint has elastic displacement data without provenance");
  }
  if (!std::isfinite(c.target_mass_code) || !std::isfinite(c.last_lambda) || !std::isfinite(c.last_residual_code) ||
      !std::isfinite(c.last_derivative_code) || !all_finite(c.phi) || !all_finite(c.Y) || !all_finite(c.xB) ||
      !all_finite(c.dY_dt_prev))
    return set_err(err, "checkpoint contains a non-finite value");
  return true;
}

// pf_zero_mode_checkpoint.cpp:284-483: V4 with an elastic warm state, V3 without
bool write(const std::string& path, Checkpoint& c, std::string* err) {
  if (!validate(c, err)) return false;
  const Provenance& p = c.prov;
  const bool v4 = c.elastic_present;
  Image h;
  h.put_text(v4 ? "PFZMCHK4" : "PFZMCHK3", 8);
  h.put<std::uint32_t>(v4 ? 4u : 3u);
  h.put<std::uint32_t>(v4 ? kHeaderV4 : kHeaderV3);
  h.put<std::uint64_t>(c.phi.size());
  if (v4) h.put<std::uint64_t>(c.displacement_k.size() / 6);
  h.put<std::uint64_t>(c.accepted_step);
  h.put<std::int32_t>(c.nx);
  h.put<std::int32_t>(c.ny);
  h.put<std::int32_t>(c.nz);
  h.put<std::int32_t>(v4 ? 1 : 0);   // V4 elastic_state_present / V3 reserved
  h.put<double>(c.dt_code);
  h.put<double>(c.temperature_K);
  h.put<double>(c.target_mass_code);
  h.put<double>(c.last_lambda);
  h.put<double>(c.last_residual_code);
  h.put<double>(c.last_derivative_code);
  h.put<std::uint64_t>(c.last_iterations);
  h.put<std::uint64_t>(c.accepted_zero_mode_steps);
  h.put<std::uint64_t>(p.parameter_fingerprint);
  if (v4) {
    h.put<std::uint64_t>(p.elastic_solver_fingerprint);
    h.put<std::uint64_t>(c.elastic_source_field_step);
    h.put<std::uint64_t>(c.elastic_last_iterations);
    h.put<double>(c.elastic_last_relative_residual);
  }
  const std::string* sel[7] = {&p.zero_mode,        &p.backend,
                               &p.composition_mode, &p.y_update_mode,
                               &p.explicit_context, &p.reaction_discretization,
                               &p.elastic_solver_mode};
  const char* sel_name[7] = {"zero_mode",        "backend",
                             "composition_mode", "y_update_mode",
                             "explicit_context", "reaction_discretization",
                             "elastic_solver_mode"};
  for (int i = 0; i < (v4 ? 7 : 6); ++i) {
    if (!text_fits(*sel[i], kSelectorBytes, sel_name[i], err)) return false;
    h.put_text(*sel[i], kSelectorBytes);
  }
  const std::string* ids[3] = {&p.initial_state_class, &p.fixture_manifest_sha256,
                               &p.profile_library_manifest_sha256};
  const char* id_name[3] = {"initial_state_class", "fixture_manifest_sha256", "profile_library_manifest_sha256"};
  for (int i = 0; i < 3; ++i) {
    if (!text_fits(*ids[i], kIdentityBytes, id_name[i], err)) return false;
    h.put_text(*ids[i], kIdentityBytes);
  }
  const size_t sum_at = h.b.size();
  h.put<std::uint64_t>(0);
  if (h.b.size() != (v4 ? kHeaderV4 : kHeaderV3)) return set_err(err, "checkpoint header layout error");

  std::uint64_t sum = fnv1a64(h.b.data(), h.b.size());
  const std::vector<double>* fields[4] = {&c.phi, &c.Y, &c.xB, &c.dY_dt_prev};
  for (const auto* f : fields) sum = fnv1a64(f->data(), f->size() * sizeof(double), sum);
  if (v4) sum = fnv1a64(c.displacement_k.data(), c.displacement_k.size() * sizeof(float), sum);
  std::memcpy(h.b.data() + sum_at, &sum, sizeof(sum));

  const std::string tmp = path + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return set_err(err, "cannot open checkpoint temporary file: " + std::string(std::strerror(errno)));
  bool ok = write_all(f, h.b.data(), h.b.size());
  for (const auto* fl : fields)
    if (ok) ok = write_all(f, fl->data(), fl->size() * sizeof(double));
  if (ok && v4) ok = write_all(f, c.displacement_k.data(), c.displacement_k.size() * sizeof(float));
  if (ok) ok = std::fflush(f) == 0;
  if (std::fclose(f) != 0) ok = false;
  if (!ok) {
    std::remove(tmp.c_str());
    return set_err(err, v4 ? "checkpoint V4 serialization failed" : "checkpoint serialization failed");
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) {
    const std::string why = std::strerror(errno);
    std::remove(tmp.c_str());
    return set_err(err, std::string(v4 ? "checkpoint V4 atomic rename failed: " : "checkpoint atomic rename failed: ") +
                            why);
  }
  c.version = v4 ? 4 : 3;
  c.payload_checksum = sum;
  return true;
}

// pf_zero_mode_checkpoint.cpp:485-760
bool read(const std::string& path, const Provenance& ex, Checkpoint* out, std::string* err) {
  if (!out) return set_err(err, "checkpoint output is null");
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return set_err(err, "cannot open checkpoint: " + std::string(std::strerror(errno)));
  unsigned char prefix[16];
  if (!read_all(f, prefix, sizeof(prefix)) || std::fseek(f, 0L, SEEK_SET) != 0) {
    std::fclose(f);
    return set_err(err, "checkpoint header/version mismatch");
  }
  std::uint32_t version, hbytes;
  std::memcpy(&version, prefix + 8, 4);
  std::memcpy(&hbytes, prefix + 12, 4);
  auto magic_is = [&](const char* m) { return std::memcmp(prefix, m, 8) == 0; };
  const bool is4 = magic_is("PFZMCHK4") && version == 4 && hbytes == kHeaderV4;
  const bool is3 = magic_is("PFZMCHK3") && version == 3 && hbytes == kHeaderV3;
  const bool is2 = magic_is("PFZMCHK2") && version == 2 && hbytes == kHeaderV2;
  if (!is4 && !is3 && !is2) {
    std::fclose(f);
    return set_err(err, "checkpoint header/version mismatch");
  }
  if (is3 && elastic_expected(ex)) {
    std::fclose(f);
    return set_err(err, "checkpoint V3 lacks required elastic runtime state");
  }
  if (is2 && (!legacy_identity(ex) || elastic_expected(ex))) {
    std::fclose(f);
    return set_err(err, "legacy checkpoint lacks required fixture/library/elastic provenance");
  }
  std::vector<unsigned char> hb(hbytes);
  bool ok = read_all(f, hb.data(), hbytes);
  Checkpoint c;
  c.version = version;
  Cursor r{hb.data() + 16};
  std::uint64_t count = 0, kcount = 0;
  std::int32_t present = 0;
  bool prov_ok = ok;
  if (ok) {
    count = r.get<std::uint64_t>();
    if (is4) kcount = r.get<std::uint64_t>();
    c.accepted_step = r.get<std::uint64_t>();
    c.nx = r.get<std::int32_t>();
    c.ny = r.get<std::int32_t>();
    c.nz = r.get<std::int32_t>();
    present = r.get<std::int32_t>();
    c.dt_code = r.get<double>();
    c.temperature_K = r.get<double>();
    c.target_mass_code = r.get<double>();
    c.last_lambda = r.get<double>();
    c.last_residual_code = r.get<double>();
    c.last_derivative_code = r.get<double>();
    c.last_iterations = r.get<std::uint64_t>();
    c.accepted_zero_mode_steps = r.get<std::uint64_t>();
    const std::uint64_t pfp = r.get<std::uint64_t>();
    std::uint64_t efp = 0;
    if (is4) {
      efp = r.get<std::uint64_t>();
      c.elastic_source_field_step = r.get<std::uint64_t>();
      c.elastic_last_iterations = r.get<std::uint64_t>();
      c.elastic_last_relative_residual = r.get<double>();
    }
    prov_ok = r.text_equals(ex.zero_mode, kSelectorBytes) && r.text_equals(ex.backend, kSelectorBytes) &&
              r.text_equals(ex.composition_mode, kSelectorBytes) && r.text_equals(ex.y_update_mode, kSelectorBytes) &&
              r.text_equals(ex.explicit_context, kSelectorBytes) &&
              r.text_equals(ex.reaction_discretization, kSelectorBytes);
    if (is4) prov_ok = prov_ok && r.text_equals(ex.elastic_solver_mode, kSelectorBytes);
    if (!is2)
      prov_ok = prov_ok && r.text_equals(ex.initial_state_class, kIdentityBytes) &&
                r.text_equals(ex.fixture_manifest_sha256, kIdentityBytes) &&
                r.text_equals(ex.profile_library_manifest_sha256, kIdentityBytes);
    prov_ok = prov_ok && pfp == ex.parameter_fingerprint;
    if (is4) prov_ok = prov_ok && present == 1 && elastic_expected(ex) && efp == ex.elastic_solver_fingerprint;
  }
  if (!ok || !prov_ok) {
    std::fclose(f);
    return set_err(err, is4 ? "checkpoint V4 elastic/zero-mode provenance mismatch"
                            : (is3 ? "checkpoint provenance mismatch" : "checkpoint zero-mode provenance mismatch"));
  }
  std::uint64_t expected_sum;
  std::memcpy(&expected_sum, hb.data() + hbytes - 8, 8);
  std::memset(hb.data() + hbytes - 8, 0, 8);
  std::uint64_t sum = fnv1a64(hb.data(), hbytes);

  if (count == 0 || count > (std::uint64_t)(std::numeric_limits<std::size_t>::max() / sizeof(double))) {
    std::fclose(f);
    return set_err(err, "checkpoint element count is invalid");
  }
  if (is4 && (kcount == 0 || kcount != k_count(c.nx, c.ny, c.nz) ||
              kcount > (std::uint64_t)(std::numeric_limits<std::size_t>::max() / (6 * sizeof(float))))) {
    std::fclose(f);
    return set_err(err, "checkpoint elastic k-space element count is invalid");
  }
  c.prov = ex;
  c.elastic_present = is4;
  std::vector<double>* fields[4] = {&c.phi, &c.Y, &c.xB, &c.dY_dt_prev};
  for (auto* fl : fields) {
    fl->resize(count);
    if (ok) ok = read_all(f, fl->data(), count * sizeof(double));
    if (ok) sum = fnv1a64(fl->data(), count * sizeof(double), sum);
  }
  if (is4) {
    c.displacement_k.resize(6 * kcount);
    if (ok) ok = read_all(f, c.displacement_k.data(), c.displacement_k.size() * sizeof(float));
    if (ok) sum = fnv1a64(c.displacement_k.data(), c.displacement_k.size() * sizeof(float), sum);
  }
  unsigned char extra;
  if (ok && std::fread(&extra, 1, 1, f) != 0) ok = false;
  std::fclose(f);
  if (!ok || sum != expected_sum) return set_err(err, "checkpoint payload checksum mismatch");
  if (!validate(c, err)) return false;
  c.payload_checksum = sum;
  *out = std::move(c);
  return true;
}

// ------------------------------------------------------------ exact-resume companion
namespace {
constexpr char kExactMagic[8] = {'P', 'F', 'A', 'M', 'D', 'X', 'S', '1'};
constexpr std::uint32_t kExactHeader = 48;    // version 1: warm state and history
constexpr std::uint32_t kExactHeader2 = 56;   // version 2: + count of the phi* block
constexpr std::uint32_t kExactHeader3 = 72;   // version 3: + the tile plan (3 int32 and one zero int32)
}  // namespace

bool write_exact(const std::string& path, const ExactState& s, std::string* err) {
  if (s.states.size() != (size_t)s.nstates * 6 * s.n) return set_err(err, "exact-resume state size mismatch");
  const bool v2 = !s.phi_star.empty();
  if (v2 && s.phi_spec.size() != 2 * s.n) return set_err(err, "exact-resume phi* spectrum size mismatch");
  Image h;
  h.put_text(std::string(kExactMagic, 8), 8);
  h.put<std::uint32_t>(3);
  h.put<std::uint32_t>(kExactHeader3);
  h.put<std::uint64_t>(s.bound_checksum);
  h.put<std::uint64_t>(s.accepted_step);
  h.put<std::int32_t>(s.precision);
  h.put<std::int32_t>(s.nstates);
  h.put<std::uint64_t>(s.n);
  h.put<std::uint64_t>(s.phi_star.size());
  for (int i = 0; i < 3; ++i) h.put<std::int32_t>(s.plan[i]);
  h.put<std::int32_t>(0);
  std::uint64_t sum = fnv1a64(h.b.data(), h.b.size());
  sum = fnv1a64(s.states.data(), s.states.size() * sizeof(double), sum);
  if (v2) {
    sum = fnv1a64(s.phi_star.data(), s.phi_star.size() * sizeof(double), sum);
    sum = fnv1a64(s.phi_spec.data(), s.phi_spec.size() * sizeof(double), sum);
  }
  const std::string tmp = path + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return set_err(err, "cannot open exact-resume temporary file: " + std::string(std::strerror(errno)));
  bool ok = write_all(f, h.b.data(), h.b.size()) && write_all(f, s.states.data(), s.states.size() * sizeof(double));
  if (v2)
    ok = ok && write_all(f, s.phi_star.data(), s.phi_star.size() * sizeof(double)) &&
         write_all(f, s.phi_spec.data(), s.phi_spec.size() * sizeof(double));
  ok = ok && write_all(f, &sum, sizeof(sum)) && std::fflush(f) == 0;
  if (std::fclose(f) != 0) ok = false;
  if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) {
    std::remove(tmp.c_str());
    return set_err(err, "exact-resume state write failed");
  }
  return true;
}

bool read_exact(const std::string& path, ExactState* s, std::string* err) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    if (err) err->clear();
    return false;
  }
  unsigned char hb[kExactHeader3];
  bool ok = read_all(f, hb, 16) && std::memcmp(hb, kExactMagic, 8) == 0;
  std::uint32_t version = 0, hsize = 0;
  if (ok) {
    Cursor r{hb + 8};
    version = r.get<std::uint32_t>();
    hsize = r.get<std::uint32_t>();
    ok = (version == 1 && hsize == kExactHeader) || (version == 2 && hsize == kExactHeader2) ||
         (version == 3 && hsize == kExactHeader3);
  }
  ok = ok && read_all(f, hb + 16, hsize - 16);
  ExactState t;
  std::uint64_t nphi = 0;
  if (ok) {
    Cursor r{hb + 16};
    t.bound_checksum = r.get<std::uint64_t>();
    t.accepted_step = r.get<std::uint64_t>();
    t.precision = r.get<std::int32_t>();
    t.nstates = r.get<std::int32_t>();
    t.n = r.get<std::uint64_t>();
    if (version >= 2) nphi = r.get<std::uint64_t>();
    if (version >= 3)
      for (int i = 0; i < 3; ++i) t.plan[i] = r.get<std::int32_t>();
    ok = t.nstates >= 0 && t.nstates <= 4 && (t.nstates > 0 || nphi > 0) && t.n > 0 && t.n < (1ull << 40) &&
         nphi < (1ull << 40);
    for (int i = 0; i < 3; ++i) ok = ok && t.plan[i] >= -1 && t.plan[i] <= 3;
  }
  std::uint64_t sum = 0, stored = 0;
  if (ok) {
    t.states.resize((size_t)t.nstates * 6 * t.n);
    t.phi_star.resize(nphi);
    t.phi_spec.resize(nphi ? 2 * t.n : 0);
    ok = read_all(f, t.states.data(), t.states.size() * sizeof(double)) &&
         read_all(f, t.phi_star.data(), t.phi_star.size() * sizeof(double)) &&
         read_all(f, t.phi_spec.data(), t.phi_spec.size() * sizeof(double)) && read_all(f, &stored, sizeof(stored));
    unsigned char extra;
    if (ok && std::fread(&extra, 1, 1, f) != 0) ok = false;
    sum = fnv1a64(t.states.data(), t.states.size() * sizeof(double), fnv1a64(hb, hsize));
    if (nphi) {
      sum = fnv1a64(t.phi_star.data(), t.phi_star.size() * sizeof(double), sum);
      sum = fnv1a64(t.phi_spec.data(), t.phi_spec.size() * sizeof(double), sum);
    }
  }
  std::fclose(f);
  if (!ok || sum != stored) return set_err(err, "exact-resume state is corrupt");
  *s = std::move(t);
  return true;
}

}  // namespace ckpt

// ---------------------------------------------------------------- fingerprints
// pf_zero_mode_parameter_fingerprint (main_cuda.cu:4187-4221): FNV-1a over the
// step's physical controls in PFParams declaration types (int grid, double
// scalars, int flags), then the NUL-terminated composition and Y-update mode
// names and the two time-level selectors.  This library runs the reference's
// default modes, "legacy" and "lagged_rhs" (main_cuda.cu:16107-16108).
std::uint64_t parameter_fingerprint(const pf_params* p) {
  std::uint64_t h = 1469598103934665603ULL;
  auto add = [&](const void* d, size_t n) { h = ckpt::fnv1a64(d, n, h); };
  const std::int32_t grid[3] = {p->Nx, p->Ny, p->Nz};
  for (int i = 0; i < 3; ++i) add(&grid[i], 4);
  const double a[] = {p->dt, p->temperature_C, p->dx, p->dy, p->dz, p->v_A, p->v_B, p->D_alpha,
                      p->D_compound, p->L_phi, p->kappa_phi, p->W, p->Y_clip, p->xB_eps};
  for (double v : a) add(&v, 8);
  const std::int32_t fl[2] = {p->enable_Y_rhs_previous_time_level, p->disable_Y_rhs_gamma_term};
  for (int i = 0; i < 2; ++i) add(&fl[i], 4);
  const char* text[4] = {"legacy", "lagged_rhs", ckpt::kExplicitContextNV1, ckpt::kReactionTangentNV1};
  for (const char* t : text) add(t, std::strlen(t) + 1);
  return h == 0 ? 1 : h;
}

// elastic_solver_parameter_fingerprint (main_cuda.cu:4223-4267)
std::uint64_t elastic_solver_fingerprint(const pf_params* p) {
  std::uint64_t h = 1469598103934665603ULL;
  auto add = [&](const void* d, size_t n) { h = ckpt::fnv1a64(d, n, h); };
  const std::int32_t grid[3] = {p->Nx, p->Ny, p->Nz};
  for (int i = 0; i < 3; ++i) add(&grid[i], 4);
  const double dxyz[3] = {p->dx, p->dy, p->dz};
  add(dxyz, sizeof(dxyz));
  const std::int32_t ctl[5] = {p->elastic_enabled, p->elastic_warm_start_enabled,
                               p->elastic_residual_control_enabled, p->elastic_iter_min, p->elastic_iter_max};
  add(ctl, sizeof(ctl));
  const double tol[2] = {p->elastic_residual_tolerance, p->elastic_residual_absolute_floor};
  add(tol, sizeof(tol));
  const std::int32_t failc = p->elastic_fail_on_nonconvergence;
  add(&failc, 4);
  double stiff[42];
  for (int i = 0; i < 21; ++i) {
    stiff[i] = p->S[i];
    stiff[21 + i] = p->S_p[i];
  }
  add(stiff, sizeof(stiff));
  double strain[13];
  for (int i = 0; i < 6; ++i) strain[i] = p->eps00[i];
  strain[6] = p->eps_iso_over_vB;
  for (int i = 0; i < 6; ++i) strain[7 + i] = p->E0[i];
  add(strain, sizeof(strain));
  return h == 0 ? 1 : h;
}

// accelerated elastic solver in the reference's sense (main_cuda.cu:24447-24450)
bool elastic_state_required(const pf_params* p) {
  return p->elastic_enabled && (p->elastic_warm_start_enabled != 0 || p->elastic_residual_control_enabled != 0);
}

}  // namespace pf

// ------------------------------------------------------------------ C ABI
namespace {

void copy_text(char* dst, size_t width, const char* src) {
  std::memset(dst, 0, width);
  std::strncpy(dst, src, width - 1);
}
std::string field(const char* s, size_t width) { return std::string(s, strnlen(s, width)); }

pf::ckpt::Provenance to_prov(const pf_checkpoint_provenance* q) {
  pf::ckpt::Provenance p;
  p.zero_mode = field(q->zero_mode, sizeof(q->zero_mode));
  p.backend = field(q->backend, sizeof(q->backend));
  p.composition_mode = field(q->composition_mode, sizeof(q->composition_mode));
  p.y_update_mode = field(q->y_update_mode, sizeof(q->y_update_mode));
  p.explicit_context = field(q->explicit_context, sizeof(q->explicit_context));
  p.reaction_discretization = field(q->reaction_discretization, sizeof(q->reaction_discretization));
  p.elastic_solver_mode = field(q->elastic_solver_mode, sizeof(q->elastic_solver_mode));
  p.initial_state_class = field(q->initial_state_class, sizeof(q->initial_state_class));
  p.fixture_manifest_sha256 = field(q->fixture_manifest_sha256, sizeof(q->fixture_manifest_sha256));
  p.profile_library_manifest_sha256 =
      field(q->profile_library_manifest_sha256, sizeof(q->profile_library_manifest_sha256));
  p.parameter_fingerprint = q->parameter_fingerprint;
  p.elastic_solver_fingerprint = q->elastic_solver_fingerprint;
  return p;
}

}  // namespace

namespace pf {
ckpt::Provenance provenance_from_abi(const pf_checkpoint_provenance* q) { return to_prov(q); }
}  // namespace pf

extern "C" unsigned long long pf_parameter_fingerprint(const pf_params* p) {
  return p ? pf::parameter_fingerprint(p) : 0ULL;
}

extern "C" unsigned long long pf_elastic_solver_fingerprint(const pf_params* p) {
  return p ? pf::elastic_solver_fingerprint(p) : 0ULL;
}

extern "C" int pf_checkpoint_provenance_default(const pf_params* p, pf_checkpoint_provenance* out) {
  if (!p || !out) return pf::fail(PF_E_ARG, "pf_checkpoint_provenance_default: null argument");
  std::memset(out, 0, sizeof(*out));
  copy_text(out->zero_mode, sizeof(out->zero_mode), pf::ckpt::kModeV1);
  copy_text(out->backend, sizeof(out->backend), pf::ckpt::kHostBackendV1);
  copy_text(out->composition_mode, sizeof(out->composition_mode), "legacy");
  copy_text(out->y_update_mode, sizeof(out->y_update_mode), "lagged_rhs");
  copy_text(out->explicit_context, sizeof(out->explicit_context), pf::ckpt::kExplicitContextNV1);
  copy_text(out->reaction_discretization, sizeof(out->reaction_discretization), pf::ckpt::kReactionTangentNV1);
  copy_text(out->initial_state_class, sizeof(out->initial_state_class), pf::ckpt::kLegacyUnspecified);
  copy_text(out->fixture_manifest_sha256, sizeof(out->fixture_manifest_sha256), pf::ckpt::kLegacyUnspecified);
  copy_text(out->profile_library_manifest_sha256, sizeof(out->profile_library_manifest_sha256),
            pf::ckpt::kLegacyUnspecified);
  const bool el = pf::elastic_state_required(p);
  copy_text(out->elastic_solver_mode, sizeof(out->elastic_solver_mode),
            el ? pf::ckpt::kElasticWarmStartResidualV1 : pf::ckpt::kElasticNotRequired);
  out->elastic_solver_fingerprint = el ? pf::elastic_solver_fingerprint(p) : 0ULL;
  out->parameter_fingerprint = pf::parameter_fingerprint(p);
  return PF_OK;
}

extern "C" int pf_checkpoint_file_write(const char* path, const pf_checkpoint_provenance* prov,
                                        const pf_checkpoint_header* hdr, const double* phi, const double* Y,
                                        const double* xB, const double* dYdt, const float* displacement_k) {
  if (!path || !prov || !hdr || !phi || !Y || !xB || !dYdt)
    return pf::fail(PF_E_ARG, "pf_checkpoint_file_write: null argument");
  if (hdr->nx <= 0 || hdr->ny <= 0 || hdr->nz <= 0)
    return pf::fail(PF_E_IO, "checkpoint grid dimensions are invalid");
  pf::ckpt::Checkpoint c;
  c.accepted_step = hdr->accepted_step;
  c.nx = hdr->nx;
  c.ny = hdr->ny;
  c.nz = hdr->nz;
  c.dt_code = hdr->dt_code;
  c.temperature_K = hdr->temperature_K;
  c.prov = to_prov(prov);
  c.target_mass_code = hdr->target_mass_code;
  c.last_lambda = hdr->last_lambda;
  c.last_residual_code = hdr->last_residual_code;
  c.last_derivative_code = hdr->last_derivative_code;
  c.last_iterations = hdr->last_iterations;
  c.accepted_zero_mode_steps = hdr->accepted_zero_mode_steps;
  c.elastic_present = hdr->elastic_present != 0;
  c.elastic_source_field_step = hdr->elastic_source_field_step;
  c.elastic_last_iterations = hdr->elastic_last_iterations;
  c.elastic_last_relative_residual = hdr->elastic_last_relative_residual;
  const size_t n = (size_t)hdr->nx * hdr->ny * hdr->nz;
  c.phi.assign(phi, phi + n);
  c.Y.assign(Y, Y + n);
  c.xB.assign(xB, xB + n);
  c.dY_dt_prev.assign(dYdt, dYdt + n);
  if (c.elastic_present) {
    if (!displacement_k) return pf::fail(PF_E_ARG, "pf_checkpoint_file_write: elastic state without data");
    const size_t kn = 6 * (size_t)hdr->nx * hdr->ny * (hdr->nz / 2 + 1);
    c.displacement_k.assign(displacement_k, displacement_k + kn);
  }
  std::string err;
  if (!pf::ckpt::write(path, c, &err)) return pf::fail(PF_E_IO, err);
  return PF_OK;
}

extern "C" int pf_checkpoint_file_read(const char* path, const pf_checkpoint_provenance* expected,
                                       pf_checkpoint_header* hdr, double* phi, double* Y, double* xB, double* dYdt,
                                       float* displacement_k) {
  if (!path || !expected || !hdr) return pf::fail(PF_E_ARG, "pf_checkpoint_file_read: null argument");
  pf::ckpt::Checkpoint c;
  std::string err;
  if (!pf::ckpt::read(path, to_prov(expected), &c, &err)) return pf::fail(PF_E_IO, err);
  std::memset(hdr, 0, sizeof(*hdr));
  hdr->version = (int)c.version;
  hdr->accepted_step = c.accepted_step;
  hdr->nx = c.nx;
  hdr->ny = c.ny;
  hdr->nz = c.nz;
  hdr->dt_code = c.dt_code;
  hdr->temperature_K = c.temperature_K;
  hdr->target_mass_code = c.target_mass_code;
  hdr->last_lambda = c.last_lambda;
  hdr->last_residual_code = c.last_residual_code;
  hdr->last_derivative_code = c.last_derivative_code;
  hdr->last_iterations = c.last_iterations;
  hdr->accepted_zero_mode_steps = c.accepted_zero_mode_steps;
  hdr->elastic_present = c.elastic_present ? 1 : 0;
  hdr->elastic_source_field_step = c.elastic_source_field_step;
  hdr->elastic_last_iterations = c.elastic_last_iterations;
  hdr->elastic_last_relative_residual = c.elastic_last_relative_residual;
  hdr->payload_checksum = c.payload_checksum;
  const size_t n = c.phi.size();
  if (phi) std::memcpy(phi, c.phi.data(), n * sizeof(double));
  if (Y) std::memcpy(Y, c.Y.data(), n * sizeof(double));
  if (xB) std::memcpy(xB, c.xB.data(), n * sizeof(double));
  if (dYdt) std::memcpy(dYdt, c.dY_dt_prev.data(), n * sizeof(double));
  if (displacement_k && c.elastic_present)
    std::memcpy(displacement_k, c.displacement_k.data(), c.displacement_k.size() * sizeof(float));
  return PF_OK;
}
