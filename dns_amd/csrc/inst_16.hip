// Engine instantiation for n = 16 (both precisions); one file per grid size
// so the kernels compile in parallel.
#include "engine.hpp"
namespace dnsk {
Impl *make_engine_16(int prec) { return make_engine<16>(prec); }
}  // namespace dnsk
