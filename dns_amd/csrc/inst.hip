// Engine instantiation for one grid size n = DNS_N (both precisions).  The
// build compiles this file once per entry of grid_sizes.h with -DDNS_N=<n>,
// so the sizes compile in parallel; without DNS_N it is empty.
#ifdef DNS_N
#include "engine.hpp"
#define DNS_CAT2(a, b) a##b
#define DNS_CAT(a, b) DNS_CAT2(a, b)
namespace dnsk {
Impl *DNS_CAT(make_engine_, DNS_N)(int prec) { return make_engine<DNS_N>(prec); }
}  // namespace dnsk
#endif
