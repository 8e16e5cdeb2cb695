// Tile-class table and persistent-grid size of the variable-size batched
// gemm.  Indexed [bucket(m)][bucket(n)], bucket 0: <=16, 1: <=32, 2: <=64,
// 3: <=128, 4: <=256, 5: larger; class 0 = 64x64 tiles, 1 = 128x128 tiles.
// Default (untuned) rule: 128x128 tiles once both m and n exceed 128.
// mojoblas_amd/tune_vbatched.py measures the alternatives on the GPU and
// rewrites this file with the winners.
#pragma once

#define VB_BUCKETS 6
#define VB_DEFAULT_WGS_PER_CU 2

static const int VB_DEFAULT_CLASS[VB_BUCKETS][VB_BUCKETS] = {
    {0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 1, 1},
    {0, 0, 0, 0, 1, 1},
};
