// Device-resident PLINK BED: SNP-major rows, each padded to a multiple of
// 16 bytes (pad bytes are zero and always masked by sample count).
#pragma once
#include "cb_internal.h"

struct cb_genotypes {
    cb_context* ctx = nullptr;
    int64_t n = 0;              // samples (.fam lines)
    int64_t m = 0;              // SNPs (.bim lines)
    int64_t bytes_per_snp = 0;  // ceil(n/4), as in the file
    int64_t pitch = 0;          // device row pitch in bytes (multiple of 16)
    cbamd::DevBuf<uint8_t> bed;
};
