// Engine instantiation for n = 96 (both precisions); one file per grid size
// so the kernels compile in parallel.
#include "engine.hpp"
namespace dnsk {
Impl *make_engine_96(int prec) { return make_engine<96>(prec); }
}  // namespace dnsk
