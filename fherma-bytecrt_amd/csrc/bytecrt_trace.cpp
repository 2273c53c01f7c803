// ROCTx-backed trace ranges (see bytecrt_trace.hpp).
#include "bytecrt_trace.hpp"

#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdlib>
#include <cstring>

namespace bcrt {

bool trace_on() {
    static const bool on = [] {
        const char* v = std::getenv("BCRT_TRACE");
        return v && std::strcmp(v, "0") != 0 && *v != '\0';
    }();
    return on;
}

void trace_push(const char* name) { roctxRangePushA(name); }
void trace_pop() { roctxRangePop(); }

uint64_t trace_start(const char* name) { return trace_on() ? uint64_t(roctxRangeStartA(name)) : 0; }

void trace_stop(uint64_t id) {
    if (trace_on()) roctxRangeStop(roctx_range_id_t(id));
}

}  // namespace bcrt
