// Mixed-precision instantiation: FP32 transforms and spectral updates, FP64
// state, reductions, zero mode and elastic residual.
#include "pf_engine.hpp"

namespace pf {
std::unique_ptr<EngineBase> make_engine_f32() { return std::make_unique<Engine<float>>(); }
}  // namespace pf
