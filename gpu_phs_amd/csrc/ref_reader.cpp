// Reader of the reference's .ref process files (header and tree blocks): the layout
// file_input.cpp:123-210 reads.  Six "key value" header lines (n_channels, NIN, NOUT,
// NTreesForest, NTreesGrove, nx), then per tree seven lines (daughters1, daughters2,
// haschildren: n_prt bincodes/flags each; hasfriends: one int, ignored as the reference
// does at file_input.cpp:172-173; mappings, map_masses, map_widths: n_prt_out each), then
// flv_masses and flv_widths lines of n_in + n_out values.  Event blocks are not read.
#include <fstream>
#include <sstream>
#include <string>

#include "phs_state.h"

using namespace gphs;

extern "C" int gpu_phs_process_load_ref(const char* path, double E1, double E2) {
  std::ifstream in(path ? path : "");
  if (!in) return set_error(GPU_PHS_E_IO, std::string("cannot open ") + (path ? path : "(null)"));
  int header[6];
  std::string line, key;
  for (int i = 0; i < 6; i++) {
    if (!std::getline(in, line)) return set_error(GPU_PHS_E_IO, "truncated header");
    std::stringstream ss(line);
    if (!(ss >> key >> header[i])) return set_error(GPU_PHS_E_IO, "bad header line: " + line);
  }
  const int n_channels = header[0], n_in = header[1], n_out = header[2], nx = header[5];
  if (n_in != 2 || n_out < 2 || n_out > kMaxOut || n_channels < 1)
    return set_error(GPU_PHS_E_IO, "unsupported header sizes");
  if (nx != 3 * n_out - 4) return set_error(GPU_PHS_E_IO, "header nx differs from 3 n_out - 4");
  const int n_prt = (1 << (n_in + n_out)) - 1, n_prt_out = (1 << n_out) - 1;
  std::vector<int> d1((size_t)n_channels * n_prt), d2(d1.size()), hc(d1.size()), mid((size_t)n_channels * n_prt_out);
  std::vector<double> mm(mid.size()), mw(mid.size()), flv(n_in + n_out), flw(n_in + n_out);
  for (int c = 0; c < n_channels; c++) {
    for (int l = 0; l < 7; l++) {
      if (!std::getline(in, line)) return set_error(GPU_PHS_E_IO, "truncated tree block");
      std::stringstream ss(line);
      ss >> key;
      bool good = true;
      switch (l) {
        case 0: for (int i = 0; i < n_prt; i++) good &= bool(ss >> d1[(size_t)c * n_prt + i]); break;
        case 1: for (int i = 0; i < n_prt; i++) good &= bool(ss >> d2[(size_t)c * n_prt + i]); break;
        case 2: for (int i = 0; i < n_prt; i++) good &= bool(ss >> hc[(size_t)c * n_prt + i]); break;
        case 3: break;
        case 4: for (int i = 0; i < n_prt_out; i++) good &= bool(ss >> mid[(size_t)c * n_prt_out + i]); break;
        case 5: for (int i = 0; i < n_prt_out; i++) good &= bool(ss >> mm[(size_t)c * n_prt_out + i]); break;
        default: for (int i = 0; i < n_prt_out; i++) good &= bool(ss >> mw[(size_t)c * n_prt_out + i]); break;
      }
      if (!good) return set_error(GPU_PHS_E_IO, "bad tree line: " + key);
    }
  }
  for (int l = 0; l < 2; l++) {
    if (!std::getline(in, line)) return set_error(GPU_PHS_E_IO, "missing flavour lines");
    std::stringstream ss(line);
    ss >> key;
    for (int i = 0; i < n_in + n_out; i++)
      if (!(ss >> (l == 0 ? flv[i] : flw[i]))) return set_error(GPU_PHS_E_IO, "bad flavour line");
  }
  return gpu_phs_process_create(n_in, n_out, n_channels, d1.data(), d2.data(), hc.data(), nullptr, mid.data(),
                                mm.data(), mw.data(), flv.data(), flw.data(), E1, E2);
}
