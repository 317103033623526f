// The device-resident request loop (dloop.hip) behind a server created with LUX_SGM_SERVER_DEVICE_LOOP.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/lux_sgm.h"

namespace lux {

struct DLoop;
// budget: the bytes of the device pool every admitted request's buffers are placed in (required, > 0).
int dloop_create(int device, uint64_t budget, int slots, DLoop** out, std::string* err);
// Stops the loop once the running requests are done (pending ones are dropped) and frees everything.
void dloop_destroy(DLoop* d);
int dloop_submit(DLoop* d, const lux_sgm_rect_params* p, int mem, const void* left, const void* right, uint16_t* disp,
                 float* sub, uint8_t* valid, uint64_t* ticket, std::string* err);
int dloop_poll(DLoop* d, uint64_t ticket, int32_t* state, std::string* err);
int dloop_wait(DLoop* d, uint64_t ticket, int32_t timeout_ms, lux_sgm_request_stats* stats, std::string* err);
int dloop_info(DLoop* d, lux_sgm_server_info* info);

}  // namespace lux
