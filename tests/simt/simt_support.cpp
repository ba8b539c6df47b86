// Error plumbing for the CPU (SIMT stand-in) build of the batched solver:
// the same thread-local last-error string and fail() as the HIP library's
// C ABI, so fsp_batch_* report errors through fsp_last_error().
#include <string>

#include "fsp_batch_impl.h"

namespace fsp {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

fsp_status fail(fsp_status st, const std::string &msg)
{
    set_error(msg);
    return st;
}

// The stream scheduler drives the HIP runtime's streams and the
// single-instance engine, which this CPU build does not contain; the
// simulated device always takes the persistent-kernel path.
fsp_status batch_solve_streams(BatchImpl &)
{
    return fail(FSP_ERR_UNSUPPORTED, "no stream scheduler in the CPU build");
}
fsp_status batch_autotune_streams(BatchImpl &, double *, int *)
{
    return fail(FSP_ERR_UNSUPPORTED, "no stream scheduler in the CPU build");
}
void batch_streams_free(BatchImpl &) {}

}  // namespace fsp

extern "C" const char *fsp_last_error(void) { return fsp::g_last_error.c_str(); }
