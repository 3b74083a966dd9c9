// Temporary entry points replaced by rtg_serve.cpp / rtg_xml.cpp.
#include "rtg_host.h"

extern "C" {
int rtg_scene_load_xml(rtg_ctx* ctx, const char*, rtg_scene**, rtg_render_desc*) {
  if (ctx) ctx->err = "xml loader not built";
  return RTG_ERR_INVALID;
}
int rtg_server_create(rtg_ctx* ctx, uint64_t, rtg_server**) {
  if (ctx) ctx->err = "server not built";
  return RTG_ERR_INVALID;
}
int rtg_server_destroy(rtg_server*) { return RTG_ERR_INVALID; }
int rtg_submit(rtg_server*, rtg_scene*, const rtg_render_desc*, int64_t*) { return RTG_ERR_INVALID; }
int rtg_server_step(rtg_server*, double) { return RTG_ERR_INVALID; }
int rtg_poll(rtg_server*, int64_t, float*, double*) { return RTG_ERR_INVALID; }
int rtg_cancel(rtg_server*, int64_t) { return RTG_ERR_INVALID; }
}
