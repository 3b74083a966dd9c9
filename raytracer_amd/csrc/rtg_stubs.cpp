// Temporary entry points replaced by rtg_xml.cpp.
#include "rtg_host.h"

extern "C" {
int rtg_scene_load_xml(rtg_ctx*, const char*, rtg_scene**, rtg_render_desc*) {
  rtg::set_error("xml loader not built");
  return RTG_ERR_INVALID;
}
}
