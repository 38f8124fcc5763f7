// Engine instantiation for n = 128 (both precisions); one file per grid size
// so the kernels compile in parallel.
#include "engine.hpp"
namespace dnsk {
Impl *make_engine_128(int prec) { return make_engine<128>(prec); }
}  // namespace dnsk
