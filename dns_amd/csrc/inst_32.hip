// Engine instantiation for n = 32 (both precisions); one file per grid size
// so the kernels compile in parallel.
#include "engine.hpp"
namespace dnsk {
Impl *make_engine_32(int prec) { return make_engine<32>(prec); }
}  // namespace dnsk
