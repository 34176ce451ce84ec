// Precision dispatch over the per-precision launch tables.
#include "registry.hpp"

namespace af {

int brick_by_f32(int nx2);
int brick_bx_f32(int ny);
const RowCfg* find_row_cfg_f32(int nx2, int bx);
const ColCfg* find_col_cfg_f32(int ny, int by);
const PrecCfg* find_prec_cfg_f32();
int brick_by_f64(int nx2);
int brick_bx_f64(int ny);
const RowCfg* find_row_cfg_f64(int nx2, int bx);
const ColCfg* find_col_cfg_f64(int ny, int by);
const PrecCfg* find_prec_cfg_f64();

int brick_by(int precision, int nx2) { return precision == 64 ? brick_by_f64(nx2) : brick_by_f32(nx2); }
int brick_bx(int precision, int ny) { return precision == 64 ? brick_bx_f64(ny) : brick_bx_f32(ny); }
const RowCfg* find_row_cfg(int precision, int nx2, int bx) {
  return precision == 64 ? find_row_cfg_f64(nx2, bx) : find_row_cfg_f32(nx2, bx);
}
const ColCfg* find_col_cfg(int precision, int ny, int by) {
  return precision == 64 ? find_col_cfg_f64(ny, by) : find_col_cfg_f32(ny, by);
}
const PrecCfg* find_prec_cfg(int precision) {
  return precision == 64 ? find_prec_cfg_f64() : find_prec_cfg_f32();
}

}  // namespace af
