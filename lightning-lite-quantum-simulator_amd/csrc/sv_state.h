// sv_state.h -- definition of the opaque llsv_state (internal to libllsv).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/llsv.h"
#include "sv_common.h"
#include "sv_program.h"

struct llsv_state {
    int n = 0;
    uint64_t size = 0;
    int device = 0;
    llsv::v2d* amps = nullptr;
    hipStream_t stream = nullptr;
    double* partials = nullptr;      // 2 * kRedBlocks doubles
    double* result_dev = nullptr;    // 4 doubles
    double* result_host = nullptr;   // 4 doubles, pinned
    unsigned char* params = nullptr; // device parameter area
    size_t params_bytes = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t marks[LLSV_MARK_SLOTS] = {};
    llsv::ProgramCache* program = nullptr;
    // llsv_probabilities: device chunk of kProbChunk doubles, allocated on first use and kept
    double* prob_chunk = nullptr;
    // llsv_apply_dense: pinned upload slots (matrix + offset table), each reused only after the
    // copy that last read it has completed (its event)
    unsigned char* stage = nullptr;
    hipEvent_t stage_ev[8] = {};
    int stage_next = 0;
};
