// Error plumbing shared by the C ABI translation units.
#pragma once
#include <exception>
#include <string>

namespace sxh {
// Thread-local last error; returns -1 so callers can `return cxh_fail(...)`.
int cxh_fail(const std::string& msg);
void cxh_clear_error();
}  // namespace sxh

#define CXH_TRY try {
#define CXH_CATCH                              \
  }                                            \
  catch (const std::exception& e) {            \
    return ::sxh::cxh_fail(e.what());          \
  }                                            \
  catch (...) {                                \
    return ::sxh::cxh_fail("unknown exception"); \
  }

using sxh::cxh_fail;
