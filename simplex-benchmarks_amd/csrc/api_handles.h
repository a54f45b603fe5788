// Opaque handle types behind the C ABI and the exception-to-status wrapper
// shared by the api*.cpp translation units.
#pragma once
#include <memory>
#include <stdexcept>

#include "mg.h"
#include "op.h"

struct mf_operator {
  smf::Operator op;
};

struct mf_multigrid {
  std::unique_ptr<smf::Multigrid> mg;
};

#define MF_TRY(body)                      \
  try {                                   \
    body;                                 \
    return 0;                             \
  } catch (const std::exception& e) {     \
    smf::set_error(e.what());             \
    return -1;                            \
  } catch (...) {                         \
    smf::set_error("unknown error");      \
    return -1;                            \
  }

inline void require_device(const smf::Operator& op) {
  if (!op.on_device) throw std::runtime_error("operator was created with host_only=1: no device data");
}
