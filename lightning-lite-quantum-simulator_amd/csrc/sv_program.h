// sv_program.h -- the gate-program path (whole circuits handed over at once).
#pragma once

#include "../../include/llsv.h"

namespace llsv {

// Host-side cache kept per state for the program scheduler.
struct ProgramCache {
    int reserved = 0;
};

int run_program(llsv_state* s, const llsv_op* ops, uint64_t n_ops);

}  // namespace llsv
