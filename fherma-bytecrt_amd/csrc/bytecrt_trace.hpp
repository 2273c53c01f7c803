// Host-side trace ranges (ROCTx) for rocprofv3 --marker-trace: whole calls,
// operand packing, waits and unpacking, server waves.  Off unless the
// environment sets BCRT_TRACE=1 (checked once per process); when off a range
// costs one predictable branch.
#pragma once

#include <cstdint>

namespace bcrt {

bool trace_on();
void trace_push(const char* name);
void trace_pop();
uint64_t trace_start(const char* name);  // a range that may end on another thread
void trace_stop(uint64_t id);

// Scoped range on the current thread.
class TraceScope {
   public:
    explicit TraceScope(const char* name) : on_(trace_on()) {
        if (on_) trace_push(name);
    }
    ~TraceScope() {
        if (on_) trace_pop();
    }
    TraceScope(const TraceScope&) = delete;
    TraceScope& operator=(const TraceScope&) = delete;

   private:
    bool on_;
};

}  // namespace bcrt
