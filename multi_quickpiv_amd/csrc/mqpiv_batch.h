// Library-internal batch entry points shared by the C ABI and the request server.
#pragma once
#include <stddef.h>

#include "../../include/mqpiv.h"

// Device workspace one request needs for its integral arrays (0 for FFTCC and
// masked NSQECC): the arrays, 256-aligned, followed by the prefix-scan temporary.
size_t request_ws_bytes(const mqpiv_request& r);

// mqpiv_piv_batch_device with each request's integral arrays built in ws[i]
// (request_ws_bytes(reqs[i]) bytes, 256-aligned) instead of the context's own
// buffers; ws == NULL is exactly mqpiv_piv_batch_device.
int batch_device_ws(mqpiv_ctx* c, const mqpiv_request* reqs, int32_t n, void* stream, void* const* ws);
