// Shared internals of the C ABI translation units.
#pragma once

#include <exception>
#include <new>
#include <string>

#include "bcx.h"
#include "bcx_host.hpp"

struct bcx_graph;
struct bcx_index;

namespace bcx {
void set_error(const std::string &m);
inline int fail(int code, const std::string &m) {
  set_error(m);
  return code;
}
const Graph &graph_of(const bcx_graph *g);
const Index &index_of(const bcx_index *ix);
}  // namespace bcx

#define BCX_GUARD_BEGIN try {
#define BCX_GUARD_END                                          \
  }                                                            \
  catch (const bcx::Error &e) {                                \
    return bcx::fail(e.code, e.msg);                           \
  }                                                            \
  catch (const std::bad_alloc &) {                             \
    return bcx::fail(BCX_ENOMEM, "host allocation failed");    \
  }                                                            \
  catch (const std::exception &e) {                            \
    return bcx::fail(BCX_EINVAL, e.what());                    \
  }
