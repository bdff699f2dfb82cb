// Internal pieces of the extern "C" boundary (include/spaqr.h) shared by the
// translation units that define its entry points: the opaque handle types
// and the exception-to-status guard.
#pragma once
#include <memory>
#include <new>
#include <string>

#include "../../include/spaqr.h"
#include "host/engine.hpp"
#include "host/krylov.hpp"

struct spaqr_tree {
  std::unique_ptr<spq::EngineBase> e;
  int device = 0;
};
struct spaqr_matrix {
  std::unique_ptr<spq::DevMatrix> d;
};

namespace spq {

// Message of the calling thread's last failure (spaqr_last_error).
std::string& last_error();

// Runs f and turns its exceptions into a status code (SPAQR_E_*), keeping
// the message for spaqr_last_error.
template <typename F>
int guard(F&& f) {
  try {
    f();
    return SPAQR_OK;
  } catch (const Error& e) {
    last_error() = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    last_error() = "host allocation failed";
    return SPAQR_E_OOM;
  } catch (const std::exception& e) {
    last_error() = e.what();
    return SPAQR_E_ARG;
  }
}

inline EngineBase& eng(spaqr_tree* t) {
  if (!t || !t->e) throw Error(kErrArg, "null tree");
  return *t->e;
}
inline const EngineBase& eng(const spaqr_tree* t) {
  if (!t || !t->e) throw Error(kErrArg, "null tree");
  return *t->e;
}

}  // namespace spq
