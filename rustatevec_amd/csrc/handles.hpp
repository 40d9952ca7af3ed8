// handles.hpp -- the opaque handle types of include/qsvh.h.
#pragma once

#include "engine.hpp"
#include "ir.hpp"
#include "program.hpp"

struct qsvh_ctx {
  qsvh::Ctx c;
};
struct qsvh_state {
  qsvh::State s;
};
struct qsvh_circuit {
  qsvh::Circuit c;
};
struct qsvh_program {
  qsvh::Program p;
};
