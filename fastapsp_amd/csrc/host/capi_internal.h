// Shared by the C wrappers of the host classes: the fapGraph handle and the
// exception -> status translation.
#pragma once

#include <new>
#include <stdexcept>

#include "fap/fap.h"
#include "fap_hip.h"

struct fap_graph {
  fap::fapGraph *g = nullptr;
};

namespace fap_hip {
void set_error(const char *fmt, ...);
}

template <typename F>
int guarded(F &&f) {
  try {
    f();
    return FAP_OK;
  } catch (const fap::FapError &e) {
    fap_hip::set_error("%s", e.what());
    return e.status();
  } catch (const std::invalid_argument &e) {
    fap_hip::set_error("%s", e.what());
    return FAP_ERR_INVALID;
  } catch (const std::bad_alloc &) {
    fap_hip::set_error("host allocation failed");
    return FAP_ERR_NOMEM;
  } catch (const std::exception &e) {
    fap_hip::set_error("%s", e.what());
    return FAP_ERR_INVALID;
  }
}
