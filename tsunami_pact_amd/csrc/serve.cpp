// serve.cpp -- continuous-batching server (placeholder until the server lands).
#include "tsu_internal.h"

extern "C" int tsu_serve_open(tsu_ctx*, const tsu_serve_cfg*, tsu_server** out) { if (out) *out = nullptr; return TSU_EINTERNAL; }
extern "C" int tsu_serve_submit(tsu_server*, const char* const*, const int32_t*, const char* const*, const int32_t*, int32_t, uint64_t*) { return TSU_EINTERNAL; }
extern "C" int tsu_serve_poll(tsu_server*, uint64_t, int32_t*, tsu_req_stats*) { return TSU_EINTERNAL; }
extern "C" int tsu_serve_wait(tsu_server*, uint64_t, int32_t*, tsu_req_stats*) { return TSU_EINTERNAL; }
extern "C" int tsu_serve_close(tsu_server*) { return TSU_EINTERNAL; }
