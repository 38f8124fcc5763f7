// Engine instantiation for n = 1024 (both precisions); one file per grid size
// so the kernels compile in parallel.
#include "engine.hpp"
namespace dnsk {
Impl *make_engine_1024(int prec) { return make_engine<1024>(prec); }
}  // namespace dnsk
