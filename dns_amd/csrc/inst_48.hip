// Engine instantiation for n = 48 (both precisions); one file per grid size
// so the kernels compile in parallel.
#include "engine.hpp"
namespace dnsk {
Impl *make_engine_48(int prec) { return make_engine<48>(prec); }
}  // namespace dnsk
