// FP64 instantiations of the stage kernels and their launch table.
// Row pass: R = radix, RB = rows per block (= brick height BY), one kernel
// per brick width BX the column side may choose.  Column pass: W = columns
// per block (= brick width BX), one kernel per brick height.  Every block has
// at most 512 lanes so the compiler may use up to 256 VGPRs per lane.
#include "kernels_inst.hpp"

namespace af {
namespace {
using T = double;
const RowCfg kRow[] = {
    {16, 4, 4, 2, 16, launch_row<T, 16, 4, 4, 2>},
    {16, 4, 4, 1, 16, launch_row<T, 16, 4, 4, 1>},
    {32, 4, 8, 2, 16, launch_row<T, 32, 4, 8, 2>},
    {32, 4, 8, 1, 16, launch_row<T, 32, 4, 8, 1>},
    {64, 4, 8, 2, 32, launch_row<T, 64, 4, 8, 2>},
    {64, 4, 8, 1, 32, launch_row<T, 64, 4, 8, 1>},
    {128, 4, 8, 2, 64, launch_row<T, 128, 4, 8, 2>},
    {128, 4, 8, 1, 64, launch_row<T, 128, 4, 8, 1>},
    {256, 4, 8, 2, 128, launch_row<T, 256, 4, 8, 2>},
    {256, 4, 8, 1, 128, launch_row<T, 256, 4, 8, 1>},
    {512, 4, 8, 2, 256, launch_row<T, 512, 4, 8, 2>},
    {512, 4, 8, 1, 256, launch_row<T, 512, 4, 8, 1>},
    {1024, 4, 8, 2, 512, launch_row<T, 1024, 4, 8, 2>},
    {1024, 4, 8, 1, 512, launch_row<T, 1024, 4, 8, 1>},
    {2048, 2, 8, 2, 512, launch_row<T, 2048, 2, 8, 2>},
    {2048, 2, 8, 1, 512, launch_row<T, 2048, 2, 8, 1>},
};
const ColCfg kCol[] = {
    {32, 2, 8, 4, 8, launch_col<T, 32, 2, 8, 4>},
    {32, 2, 8, 2, 8, launch_col<T, 32, 2, 8, 2>},
    {64, 2, 8, 4, 16, launch_col<T, 64, 2, 8, 4>},
    {64, 2, 8, 2, 16, launch_col<T, 64, 2, 8, 2>},
    {128, 2, 8, 4, 32, launch_col<T, 128, 2, 8, 4>},
    {128, 2, 8, 2, 32, launch_col<T, 128, 2, 8, 2>},
    {256, 2, 8, 4, 64, launch_col<T, 256, 2, 8, 4>},
    {256, 2, 8, 2, 64, launch_col<T, 256, 2, 8, 2>},
    {512, 2, 8, 4, 128, launch_col<T, 512, 2, 8, 4>},
    {512, 2, 8, 2, 128, launch_col<T, 512, 2, 8, 2>},
    {1024, 2, 8, 4, 256, launch_col<T, 1024, 2, 8, 4>},
    {1024, 2, 8, 2, 256, launch_col<T, 1024, 2, 8, 2>},
    {2048, 2, 8, 4, 512, launch_col<T, 2048, 2, 8, 4>},
    {2048, 2, 8, 2, 512, launch_col<T, 2048, 2, 8, 2>},
    {4096, 1, 8, 4, 512, launch_col<T, 4096, 1, 8, 4>},
    {4096, 1, 8, 2, 512, launch_col<T, 4096, 1, 8, 2>},
};
const PrecCfg kPrec = {sizeof(cx<T>), launch_to_bricks<T>, launch_from_bricks<T>};
}  // namespace

int brick_by_f64(int nx2) { return nx2 >= 2048 ? 2 : 4; }
int brick_bx_f64(int ny) { return ny >= 4096 ? 1 : 2; }
const RowCfg* find_row_cfg_f64(int nx2, int bx) {
  for (const RowCfg& c : kRow)
    if (c.nx2 == nx2 && c.bx == bx) return &c;
  return nullptr;
}
const ColCfg* find_col_cfg_f64(int ny, int by) {
  for (const ColCfg& c : kCol)
    if (c.ny == ny && c.by == by) return &c;
  return nullptr;
}
const PrecCfg* find_prec_cfg_f64() { return &kPrec; }

}  // namespace af
