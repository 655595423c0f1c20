// FP64 instantiation of the time-step engine.
#include "pf_engine.hpp"

namespace pf {
std::unique_ptr<EngineBase> make_engine_f64() { return std::make_unique<Engine<double>>(); }
}  // namespace pf
