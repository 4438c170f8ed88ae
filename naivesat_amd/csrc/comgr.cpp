// In-process assembly of gfx950 code objects (amd_comgr): the native kernel templates (once
// per template and slot size) and bodies outside the encoded subset. Host code only.
#include <amd_comgr/amd_comgr.h>

#include <cstdlib>
#include <string>
#include <vector>

#include "runtime.h"

namespace ns {

// Every generated kernel is unique and the engine keeps its own in-memory cache, so comgr's
// on-disk compilation cache only adds file writes (and slows down as it fills). It is
// turned off when the library loads unless the caller has set AMD_COMGR_CACHE.
static const int kComgrCacheOff = setenv("AMD_COMGR_CACHE", "0", 0);

namespace {
struct ComgrGuard {
  std::vector<amd_comgr_data_set_t> sets;
  std::vector<amd_comgr_data_t> datas;
  amd_comgr_action_info_t info{0};
  bool have_info = false;
  ~ComgrGuard() {
    for (auto d : datas) amd_comgr_release_data(d);
    for (auto s : sets) amd_comgr_destroy_data_set(s);
    if (have_info) amd_comgr_destroy_action_info(info);
  }
};

void check_comgr(amd_comgr_status_t s, const char* what) {
  if (s == AMD_COMGR_STATUS_SUCCESS) return;
  const char* msg = "unknown";
  amd_comgr_status_string(s, &msg);
  throw Error(ST_COMPILE, std::string(what) + ": " + msg);
}

std::string comgr_log(amd_comgr_data_set_t set) {
  size_t n = 0;
  if (amd_comgr_action_data_count(set, AMD_COMGR_DATA_KIND_LOG, &n) != AMD_COMGR_STATUS_SUCCESS || !n) return "";
  amd_comgr_data_t d;
  if (amd_comgr_action_data_get_data(set, AMD_COMGR_DATA_KIND_LOG, 0, &d) != AMD_COMGR_STATUS_SUCCESS) return "";
  size_t sz = 0;
  amd_comgr_get_data(d, &sz, nullptr);
  std::string log(sz, '\0');
  if (sz) amd_comgr_get_data(d, &sz, &log[0]);
  amd_comgr_release_data(d);
  return log;
}
}  // namespace

std::vector<char> assemble_native(const std::string& asm_text) {
  ComgrGuard g;
  amd_comgr_data_t src;
  check_comgr(amd_comgr_create_data(AMD_COMGR_DATA_KIND_SOURCE, &src), "comgr data");
  g.datas.push_back(src);
  check_comgr(amd_comgr_set_data(src, asm_text.size(), asm_text.data()), "comgr set data");
  check_comgr(amd_comgr_set_data_name(src, "ns_kernel.s"), "comgr data name");
  amd_comgr_data_set_t in, rel, exe;
  check_comgr(amd_comgr_create_data_set(&in), "comgr set");
  g.sets.push_back(in);
  check_comgr(amd_comgr_create_data_set(&rel), "comgr set");
  g.sets.push_back(rel);
  check_comgr(amd_comgr_create_data_set(&exe), "comgr set");
  g.sets.push_back(exe);
  check_comgr(amd_comgr_data_set_add(in, src), "comgr add");
  check_comgr(amd_comgr_create_action_info(&g.info), "comgr action info");
  g.have_info = true;
  check_comgr(amd_comgr_action_info_set_isa_name(g.info, "amdgcn-amd-amdhsa--gfx950"), "comgr isa");
  check_comgr(amd_comgr_action_info_set_logging(g.info, true), "comgr logging");
  if (amd_comgr_do_action(AMD_COMGR_ACTION_ASSEMBLE_SOURCE_TO_RELOCATABLE, g.info, in, rel) !=
      AMD_COMGR_STATUS_SUCCESS)
    throw Error(ST_COMPILE, "assembly of the native kernel failed: " + comgr_log(rel).substr(0, 4000));
  if (amd_comgr_do_action(AMD_COMGR_ACTION_LINK_RELOCATABLE_TO_EXECUTABLE, g.info, rel, exe) !=
      AMD_COMGR_STATUS_SUCCESS)
    throw Error(ST_COMPILE, "link of the native kernel failed: " + comgr_log(exe).substr(0, 4000));
  amd_comgr_data_t ed;
  check_comgr(amd_comgr_action_data_get_data(exe, AMD_COMGR_DATA_KIND_EXECUTABLE, 0, &ed), "comgr executable");
  g.datas.push_back(ed);
  size_t sz = 0;
  check_comgr(amd_comgr_get_data(ed, &sz, nullptr), "comgr size");
  std::vector<char> code(sz);
  check_comgr(amd_comgr_get_data(ed, &sz, code.data()), "comgr bytes");
  return code;
}

}  // namespace ns
