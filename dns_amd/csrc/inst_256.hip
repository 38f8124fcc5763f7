// Engine instantiation for n = 256 (both precisions); one file per grid size
// so the kernels compile in parallel.
#include "engine.hpp"
namespace dnsk {
Impl *make_engine_256(int prec) { return make_engine<256>(prec); }
}  // namespace dnsk
