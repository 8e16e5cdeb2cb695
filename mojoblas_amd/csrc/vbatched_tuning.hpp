// Tile-class table and persistent-grid size of the variable-size batched
// gemm, written by mojoblas_amd/tune_vbatched.py from measurements on an
// MI355X.  Indexed [bucket(m)][bucket(n)], bucket 0: <=16, 1: <=32, 2: <=64,
// 3: <=128, 4: <=256, 5: larger; class 0 = 64x64 tiles, 1 = 128x128 tiles.
#pragma once

#define VB_BUCKETS 6

// [0] = MBH_PREC_FP32, [1] = MBH_PREC_BF16
static const int VB_DEFAULT_WGS_PER_CU[2] = {8, 6};

static const int VB_DEFAULT_CLASS[2][VB_BUCKETS][VB_BUCKETS] = {{  // FP32
    {0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0},
    {0, 0, 0, 1, 1, 1},
    {0, 0, 0, 1, 1, 1},
    {0, 0, 0, 1, 1, 1},
}, {  // BF16
    {0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0},
    {0, 0, 0, 1, 1, 1},
    {0, 0, 0, 1, 1, 1},
    {0, 0, 0, 1, 1, 1},
}};
