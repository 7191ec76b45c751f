// trace.hip -- roctx ranges for the reference's tracing spans.
//
// The reference instruments its prover with `tracing` spans ("air: prove",
// "zerocheck", "inner sumchecks", "pow grinding", "inner sum evaluation",
// "sumcheck_round", "air table: verify", ...: crates/air/src/prove.rs:48,114,133,
// 147,190; crates/sumcheck/src/prove.rs:95; crates/air/src/verify.rs:54).  Here
// the same names are roctx ranges on the host timeline, which rocprofv3
// --marker-trace records beside the kernels.  The roctx library is opened at the
// first span; without it (or with WHIRLAWAY_ROCTX=0) spans cost a null check.
#include <dlfcn.h>
#include <stdlib.h>
#include <mutex>
#include "trace.h"

namespace wl {
namespace {

using PushFn = int (*)(const char *);
using PopFn = int (*)();
PushFn g_push = nullptr;
PopFn g_pop = nullptr;
std::once_flag g_once;

void load() {
    const char *e = getenv("WHIRLAWAY_ROCTX");
    if (e && e[0] == '0') return;
    for (const char *name : {"librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so", "libroctx64.so.4",
                             "libroctx64.so"}) {
        void *h = dlopen(name, RTLD_NOW | RTLD_GLOBAL);
        if (!h) continue;
        auto push = (PushFn)dlsym(h, "roctxRangePushA");
        auto pop = (PopFn)dlsym(h, "roctxRangePop");
        if (push && pop) {
            g_pop = pop;
            g_push = push;
            return;
        }
    }
}

}  // namespace

void span_push(const char *name) {
    std::call_once(g_once, load);
    if (g_push) g_push(name);
}
void span_pop() {
    if (g_pop) g_pop();
}

}  // namespace wl
