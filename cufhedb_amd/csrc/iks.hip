// iks.hip -- batched identity key switching and small element-wise kernels.
//
// IdentityKeySwitch (thirdparty/TFHEpp/src/keyswitch.cpp:12-48):
//   b'   = b                         (same torus width)
//        = (b + 2^(W-33)) >> (W-32)  (64-bit domain -> 32-bit target)
//   for i < dn, j < t:  a_ij = ((a_i + 2^(W-1-t*basebit)) >> (W-(j+1)basebit)) & mask
//                       if a_ij != 0: res -= ksk[i][j][a_ij - 1]
// The reference runs this on the host CPU before every bootstrap
// (src/cuHEDB/gatebootstrap_gpu.cuh:42,59,...).  Here a batch is one launch
// of one of two kernels, chosen by the base: k_iks_rows (kinds 10 and 20,
// basebit 10) reads every selected key row whole, one workgroup per
// ciphertext; k_iks_tiled (kind 21, basebit 3) switches a tile of
// ciphertexts on one 64-column slice so that the few candidate rows of a
// digit position are read once per tile.  The key stays in HBM in the
// reference's own layout [dn][t][2^basebit - 1][tn + 1].
#include "launch.hpp"
#include "params.hpp"

namespace fhe {

namespace {

// Whole-row form for large bases (kinds 10 and 20: basebit 10, 1023
// candidate rows per digit position, so ciphertexts rarely share a row): one
// workgroup per ciphertext, wave w takes every fourth digit position and
// reads each selected row whole (tn + 1 words as NCH = ceil((tn + 1) / 64)
// coalesced loads, one per lane and 64 columns), two rows at a time, so the
// memory system sees 2.5 KB contiguous requests with 2 NCH loads in flight
// per wave; the four waves' partial sums meet in LDS.  The loop body has no
// branch: a zero digit reads a valid row and discards it, and lanes past the
// last column read up to 63 words beyond the row (the device key carries
// kKskPadWords spare words) into sums that are never written.
template <class TD, int DN, int T, int BB, int NCH>
__global__ __launch_bounds__(256) void k_iks_rows(const TD *__restrict__ in,
                                                  const uint32_t *__restrict__ ksk,
                                                  uint32_t *__restrict__ out, int tn, int B)
{
    constexpr int W = int(sizeof(TD) * 8);
    constexpr int NDIG = DN * T;
    static_assert(NDIG % 64 == 0, "digit positions are split in multiples of 8 pairs");
    constexpr uint32_t mask = (1u << BB) - 1;
    constexpr size_t nd = (size_t(1) << BB) - 1;
    __shared__ uint16_t dig[NDIG];
    __shared__ uint32_t part[4][NCH * 64];
    const int b = blockIdx.x;
    if (b >= B) return;
    // split gridDim.y ways over the digit positions: this workgroup takes
    // [e0, e0 + per) and adds its sum into the zero-filled output
    const int nsplit = gridDim.y, per = NDIG / nsplit, e0 = blockIdx.y * per;
    const TD *c = in + size_t(b) * (DN + 1);
    const TD prec = TD(1) << (W - (1 + BB * T));
    for (int e = e0 + threadIdx.x; e < e0 + per; e += blockDim.x) {
        const int i = e / T, j = e % T;
        dig[e] = uint16_t(((c[i] + prec) >> (W - (j + 1) * BB)) & mask);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const size_t row = size_t(tn) + 1;
    uint32_t acc[NCH];
#pragma unroll
    for (int j = 0; j < NCH; j++) acc[j] = 0;
    for (int e = e0 + wave; e < e0 + per; e += 8) {
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(uint32_t(dig[e]));
        const uint32_t d1 = __builtin_amdgcn_readfirstlane(uint32_t(dig[e + 4]));
        const uint32_t keep0 = d0 ? 0xffffffffu : 0u, keep1 = d1 ? 0xffffffffu : 0u;
        const uint32_t *k0 = ksk + (size_t(e) * nd + (d0 ? d0 - 1 : 0)) * row + lane;
        const uint32_t *k1 = ksk + (size_t(e + 4) * nd + (d1 ? d1 - 1 : 0)) * row + lane;
        uint32_t v0[NCH], v1[NCH];
#pragma unroll
        for (int j = 0; j < NCH; j++) v0[j] = k0[64 * j];
#pragma unroll
        for (int j = 0; j < NCH; j++) v1[j] = k1[64 * j];
#pragma unroll
        for (int j = 0; j < NCH; j++) acc[j] -= (v0[j] & keep0) + (v1[j] & keep1);
    }
#pragma unroll
    for (int j = 0; j < NCH; j++) part[wave][lane + 64 * j] = acc[j];
    __syncthreads();
    uint32_t body;
    if constexpr (W == 32)
        body = uint32_t(c[DN]);
    else
        body = uint32_t((c[DN] + (TD(1) << (W - 33))) >> (W - 32));
    uint32_t *o = out + size_t(b) * row;
    if (nsplit == 1) {
        for (int k = threadIdx.x; k < int(row); k += blockDim.x)
            o[k] = (k == tn ? body : 0u) + part[0][k] + part[1][k] + part[2][k] + part[3][k];
    } else {
        // integer sums mod 2^32: the order of the workgroups' additions does
        // not change the result
        for (int k = threadIdx.x; k < int(row); k += blockDim.x)
            atomicAdd(o + k, (k == tn && blockIdx.y == 0 ? body : 0u) + part[0][k] + part[1][k] +
                                 part[2][k] + part[3][k]);
    }
}

__global__ void k_fill_zero32(uint32_t *p, size_t n)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// launch k_iks_rows with NCH = ceil((tn + 1) / 64), tn <= kMaxN0, split
// ways over the digit positions (1, 2, 4 or 8; a split output is zeroed by a
// kernel first, so the launch stays capturable)
template <class TD, int DN, int T, int BB, int NCH = 1>
void launch_iks_rows(const TD *in, const uint32_t *ksk, uint32_t *out, int tn, int B, int split,
                     hipStream_t s)
{
    if constexpr (NCH < (kMaxN0 + 1 + 63) / 64) {
        if ((tn + 1 + 63) / 64 > NCH)
            return launch_iks_rows<TD, DN, T, BB, NCH + 1>(in, ksk, out, tn, B, split, s);
    }
    if (split > 1) {
        const size_t n = size_t(B) * (tn + 1);
        hipLaunchKernelGGL(k_fill_zero32, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, out, n);
    }
    hipLaunchKernelGGL((k_iks_rows<TD, DN, T, BB, NCH>), dim3(B, split), dim3(256), 0, s, in, ksk,
                       out, tn, B);
}

// Tiled form for small bases (lvl21: basebit 3, 7 key rows per digit
// position): a workgroup switches C ciphertexts on one 64-column slice of the
// output, so the 7 candidate rows of a position are read by the C
// ciphertexts from L1/L2 instead of once per ciphertext from HBM.  Each wave
// takes C/4 ciphertexts (wave-uniform, so their a_i and digits are scalar)
// and each lane one column.  The body is branch-free so that the T x C/4
// row reads of one a_i are all in flight together: a zero digit reads a
// valid row and discards it, lanes past the last column read column tn.
template <class TD, int DN, int T, int BB, int C, int TJ>
__global__ __launch_bounds__(256) void k_iks_tiled(const TD *__restrict__ in,
                                                   const uint32_t *__restrict__ ksk,
                                                   uint32_t *__restrict__ out, int tn, int B)
{
    constexpr int W = int(sizeof(TD) * 8);
    constexpr uint32_t mask = (1u << BB) - 1;
    constexpr size_t nd = (size_t(1) << BB) - 1;
    constexpr int PER = C / 4;
    const int col = blockIdx.y * 64 + (threadIdx.x & 63);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int b0 = blockIdx.x * C + wv * PER;
    const size_t row = size_t(tn) + 1;
    const TD prec = TD(1) << (W - (1 + BB * T));
    uint32_t acc[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) acc[k] = 0;
    const uint32_t *kc = ksk + min(col, tn);
    for (int i = 0; i < DN; i++) {
        TD abar[PER];
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const TD a = (b0 + k < B) ? in[size_t(b0 + k) * (DN + 1) + i] : TD(0);
            abar[k] = TD(a + prec);
        }
        // TJ x PER row reads of this a_i issued together, then the
        // subtractions
#pragma unroll
        for (int j0 = 0; j0 < T; j0 += TJ) {
            uint32_t v[TJ][PER];
#pragma unroll
            for (int jj = 0; jj < TJ; jj++) {
                const int j = j0 + jj;
                if (j >= T) break;
                const uint32_t *kr = kc + (size_t(i) * T + j) * nd * row;
#pragma unroll
                for (int k = 0; k < PER; k++) {
                    const uint32_t d = uint32_t(abar[k] >> (W - (j + 1) * BB)) & mask;
                    v[jj][k] = kr[(d ? d - 1 : 0) * row];
                }
            }
#pragma unroll
            for (int jj = 0; jj < TJ; jj++) {
                const int j = j0 + jj;
                if (j >= T) break;
#pragma unroll
                for (int k = 0; k < PER; k++) {
                    const uint32_t d = uint32_t(abar[k] >> (W - (j + 1) * BB)) & mask;
                    acc[k] -= d ? v[jj][k] : 0u;
                }
            }
        }
    }
    if (col <= tn) {
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int b = b0 + k;
            if (b >= B) continue;
            uint32_t body = 0;
            if (col == tn) {
                const TD v = in[size_t(b) * (DN + 1) + DN];
                body = W == 32 ? uint32_t(v) : uint32_t((v + (TD(1) << (W - 33))) >> (W - 32));
            }
            out[size_t(b) * row + col] = body + acc[k];
        }
    }
}

// Ciphertext-per-lane form for the lvl2 -> lvl1 switch (basebit 3, 10 digit
// positions).  A wave switches 64 ciphertexts (one per lane) on CW output
// columns; a workgroup's four waves take the same 64 ciphertexts on
// adjacent column ranges (a slice of 4 CW columns, the body column TN
// apart).  A mask word's 10 digit positions are taken in two stages of
// H = 5: the stage's 35 key rows (7 per position, contiguous in the key)
// are staged into LDS on the slice's columns, each position's rows behind a
// zero row (slot 0, the row of a zero digit), double-buffered, the next
// stage's rows in flight in registers while this stage's are used.  Each
// lane turns its own digits into LDS row offsets (two VALU per position),
// then reads its CW words of each selected row with 16-byte reads at
// immediate offsets: the selection costs no scalar work and no per-read
// address arithmetic.  Rows are padded by 16 bytes so that the (at most 8)
// distinct rows a position's 64 lanes select fall on distinct banks.  Wave
// 0 of the first slice's workgroups also switches the body column from a
// 40-word LDS copy of the rows' column TN.  The kernel is bound by its LDS
// instruction rate (about 30 LDS instructions per wave and stage).
template <class TD, int DN, int TN, int T, int BB, int CW>
__global__ __launch_bounds__(256) void k_iks_ct(const TD *__restrict__ in,
                                                const uint32_t *__restrict__ ksk,
                                                uint32_t *__restrict__ out, int B)
{
    constexpr int W = int(sizeof(TD) * 8);
    constexpr int ND = (1 << BB) - 1;  // candidate rows per digit position
    constexpr int S = ND + 1;          // LDS rows per position (slot 0: zeros)
    constexpr int H = 5;               // digit positions per stage
    constexpr int NS = T / H;          // stages per mask word
    constexpr int NG = DN * NS;        // stages
    constexpr int SR = H * ND;         // key rows per stage
    constexpr int SC = 4 * CW;         // slice columns
    constexpr int RW = SC + 4;         // LDS row words (16 bytes of padding)
    constexpr size_t row = size_t(TN) + 1;
    static_assert(T % H == 0 && H == 5 && NS == 2, "two stages of five positions");
    static_assert(CW % 8 == 0 && TN % SC == 0 && (BB * T) <= 32, "shape");
    __shared__ __align__(16) uint32_t cand[2][H * S][RW];
    __shared__ uint32_t colk[2][H * S];  // column TN of the stage's rows
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c0 = blockIdx.y * SC;
    const bool body_wave = blockIdx.y == 0 && wv == 0;
    const int b = blockIdx.x * 64 + lane;  // this lane's ciphertext
    const TD *arow = in + size_t(min(b, B - 1)) * (DN + 1);
    const TD prec = TD(1) << (W - (1 + BB * T));
    for (int e = tid; e < 2 * H * RW; e += 256) cand[e / (H * RW)][(e / RW) % H * S][e % RW] = 0u;
    if (tid < 2 * H) colk[tid / H][(tid % H) * S] = 0u;
    // stage g holds key rows g SR + q (q < SR); thread tid fetches the
    // 16-byte groups e = tid + 256 m of the SR x SC block (row e / SC4,
    // columns 4 (e % SC4) ..+3) through a buffer descriptor: a lane offset
    // fixed for the kernel plus scalar stage and row offsets, no per-load
    // address arithmetic (the key of this switch is 587,776,000 bytes,
    // inside a descriptor's 32-bit range; multi-dword buffer loads need only
    // 4-byte alignment, the key rows are 4100 bytes apart)
    constexpr int SC4 = SC / 4;
    constexpr int EL4 = SR * SC4;  // 16-byte groups per stage and slice
    constexpr int RPT = (EL4 + 255) / 256;
    static_assert(256 % SC4 == 0, "a fetch round covers whole rows");
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    constexpr uint32_t key_bytes = uint32_t(DN) * T * ND * uint32_t(row) * 4u;
    const __amdgpu_buffer_rsrc_t kd = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint32_t *>(ksk), 0, int(key_bytes), 0x00020000);
    const uint32_t vo = (uint32_t(tid / SC4) * uint32_t(row) + uint32_t(c0 + 4 * (tid % SC4))) * 4u;
    const uint32_t vob = (uint32_t(min(tid, SR - 1)) * uint32_t(row) + TN) * 4u;
    auto fetch = [&](int g, u32x4 *v, uint32_t &vb) {
        const uint32_t so = uint32_t(g) * SR * uint32_t(row) * 4u;
#pragma unroll
        for (int m = 0; m < RPT; m++) {
            const int e = tid + 256 * m;
            constexpr uint32_t step = uint32_t(256 / SC4) * uint32_t(row) * 4u;
            if (m + 1 < RPT || e < EL4)
                v[m] = __builtin_amdgcn_raw_buffer_load_b128(kd, vo, so + m * step, 0);
        }
        if (tid < SR) vb = __builtin_amdgcn_raw_buffer_load_b32(kd, vob, so, 0);
    };
    auto stage = [&](int buf, const u32x4 *v, uint32_t vb) {
#pragma unroll
        for (int m = 0; m < RPT; m++) {
            const int e = tid + 256 * m;
            const int q = e / SC4;
            if (m + 1 < RPT || e < EL4) {
                const u32x4 x = v[m];
                *reinterpret_cast<uint4 *>(&cand[buf][(q / ND) * S + q % ND + 1][4 * (e % SC4)]) =
                    uint4{x[0], x[1], x[2], x[3]};
            }
        }
        if (tid < SR) colk[buf][(tid / ND) * S + tid % ND + 1] = vb;
    };
    uint32_t acc[CW];
#pragma unroll
    for (int c = 0; c < CW; c++) acc[c] = 0;
    uint32_t accb = 0;  // body column (first slice, wave 0)
    auto consume = [&](uint32_t dig, int h, int buf) {
        const uint32_t sd = h ? dig : dig >> (BB * H);  // this stage's digits, low H * BB bits
        const uint32_t *base = &cand[buf][0][wv * CW];
        uint32_t off[H];
#pragma unroll
        for (int j = 0; j < H; j++) {
            const uint32_t d = (sd >> (BB * (H - 1 - j))) & uint32_t(ND);
            off[j] = (uint32_t(j * S) + d) * RW;
        }
#pragma unroll
        for (int c8 = 0; c8 < CW; c8 += 8) {
            uint4 x[H][2];
#pragma unroll
            for (int j = 0; j < H; j++) {
                const uint4 *p = reinterpret_cast<const uint4 *>(base + off[j] + c8);
                x[j][0] = p[0];
                x[j][1] = p[1];
            }
#pragma unroll
            for (int u = 0; u < 2; u++) {
                acc[c8 + 4 * u + 0] -= x[0][u].x + x[1][u].x + x[2][u].x + x[3][u].x + x[4][u].x;
                acc[c8 + 4 * u + 1] -= x[0][u].y + x[1][u].y + x[2][u].y + x[3][u].y + x[4][u].y;
                acc[c8 + 4 * u + 2] -= x[0][u].z + x[1][u].z + x[2][u].z + x[3][u].z + x[4][u].z;
                acc[c8 + 4 * u + 3] -= x[0][u].w + x[1][u].w + x[2][u].w + x[3][u].w + x[4][u].w;
            }
            // sums done here: at most one chunk's reads in flight
#pragma unroll
            for (int c = c8; c < c8 + 8; c++) asm volatile("" : "+v"(acc[c]));
        }
        if (body_wave) {
#pragma unroll
            for (int j = 0; j < H; j++) accb -= colk[buf][off[j] / RW];
        }
    };
    // mask words are loaded in blocks of AB, a block ahead of their use
    constexpr int AB = 4;
    static_assert(DN % AB == 0, "blocking");
    u32x4 v[RPT];
    uint32_t vb = 0;
    TD ac[AB], an[AB];
#pragma unroll
    for (int w = 0; w < AB; w++) ac[w] = arow[w];
    fetch(0, v, vb);
    stage(0, v, vb);
    __syncthreads();
    for (int w0 = 0; w0 < DN; w0 += AB) {
        if (w0 + AB < DN) {
#pragma unroll
            for (int w = 0; w < AB; w++) an[w] = arow[w0 + AB + w];
        }
#pragma unroll
        for (int u = 0; u < AB * NS; u++) {
            const int g = w0 * NS + u, h = u & 1;
            const uint32_t dig = uint32_t((ac[u / NS] + prec) >> (W - T * BB));
            if (g + 1 < NG) fetch(g + 1, v, vb);
            consume(dig, h, h);
            // keep the staging (which waits for the key rows) after the
            // subtractions, so the rows' latency hides behind them
#pragma unroll
            for (int c = 0; c < CW; c++) asm volatile("" : "+v"(acc[c]));
            if (g + 1 < NG) stage(h ^ 1, v, vb);
            __syncthreads();
        }
#pragma unroll
        for (int w = 0; w < AB; w++) ac[w] = an[w];
    }
    if (b < B) {
        uint32_t *o = out + size_t(b) * row + c0 + wv * CW;
#pragma unroll
        for (int c = 0; c < CW; c++) o[c] = acc[c];
        if (body_wave) {
            const TD x = in[size_t(b) * (DN + 1) + DN];
            const uint32_t body = W == 32 ? uint32_t(x) : uint32_t((x + (TD(1) << (W - 33))) >> (W - 32));
            out[size_t(b) * row + TN] = body + accb;
        }
    }
}

__global__ void k_tlwe1_lincomb(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b,
                                int sa, int sb, uint32_t offset, uint32_t *__restrict__ out,
                                int B)
{
    const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    constexpr int W = Lvl1::N + 1;
    if (gid >= size_t(B) * W) return;
    uint32_t v = uint32_t(sa) * a[gid] + uint32_t(sb) * b[gid];
    if (gid % W == Lvl1::N) v += offset;
    out[gid] = v;
}

__global__ void k_tlwe1_add_body(uint32_t *__restrict__ tlwe, uint32_t add, int B)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B) tlwe[size_t(b) * (Lvl1::N + 1) + Lvl1::N] += add;
}

__global__ void k_tv1_const(uint32_t *tv, uint32_t u)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < Lvl1::N) tv[i] = u;
}

// gpolygen (src/HEDB/comparison/tfhepp_utils.h:113-120): 2^scale * (i >> 6)
__global__ void k_tv1_ide(uint32_t *tv, uint32_t scale_bits)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < Lvl1::N) tv[i] = (1u << scale_bits) * uint32_t(i >> 6);
}

// out = in with `add` added to the body word (W words per TLWE)
template <class T>
__global__ void k_tlwe_copy_add_body(const T *__restrict__ in, T *__restrict__ out, int W, T add,
                                     int B)
{
    const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= size_t(B) * W) return;
    T v = in[gid];
    if (gid % W == size_t(W - 1)) v += add;
    out[gid] = v;
}

// out = sa*a + sb*b (+ add on the body word); W words per TLWE.  b == nullptr
// drops the second term; b_bcast makes every record use b's first TLWE.
template <class T>
__global__ void k_tlwe_lincomb(const T *__restrict__ a, const T *__restrict__ b, int sa, int sb,
                               T add, T *__restrict__ out, int W, int b_bcast, int B)
{
    const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= size_t(B) * W) return;
    const int e = int(gid % W);
    T v = T(sa) * a[gid];
    if (b) v += T(sb) * b[b_bcast ? size_t(e) : gid];
    if (e == W - 1) v += add;
    out[gid] = v;
}

// out = in << shift on every word (HomMSB's digit shift, msb_gpu.cuh:15-18)
template <class T>
__global__ void k_tlwe_shl(const T *__restrict__ in, T *__restrict__ out, int shift, size_t n)
{
    const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid < n) out[gid] = in[gid] << shift;
}

template <class T>
__global__ void k_tlwe_add_body(T *__restrict__ tlwe, int W, T add, int B)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B) tlwe[size_t(b) * W + W - 1] += add;
}

template <class T>
__global__ void k_tv_const(T *tv, T u, int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) tv[i] = u;
}

// gpolygen with padding 6: 2^scale * (i >> 6)
template <class T>
__global__ void k_tv_ide(T *tv, uint32_t scale_bits, int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) tv[i] = (T(1) << scale_bits) * T(i >> 6);
}

}  // namespace

template <class T>
hipError_t launch_tlwe_copy_add_body(const T *in, T *out, int W, T add, int B, hipStream_t s)
{
    if (B <= 0) return hipSuccess;
    const size_t n = size_t(B) * W;
    hipLaunchKernelGGL(k_tlwe_copy_add_body<T>, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s,
                       in, out, W, add, B);
    return hipGetLastError();
}
template hipError_t launch_tlwe_copy_add_body<uint32_t>(const uint32_t *, uint32_t *, int, uint32_t,
                                                        int, hipStream_t);
template hipError_t launch_tlwe_copy_add_body<uint64_t>(const uint64_t *, uint64_t *, int, uint64_t,
                                                        int, hipStream_t);

template <class T>
hipError_t launch_tlwe_lincomb(const T *a, const T *b, int sa, int sb, T add, T *out, int W,
                               bool b_bcast, int B, hipStream_t s)
{
    if (B <= 0) return hipSuccess;
    const size_t n = size_t(B) * W;
    hipLaunchKernelGGL(k_tlwe_lincomb<T>, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, a, b,
                       sa, sb, add, out, W, int(b_bcast), B);
    return hipGetLastError();
}
template hipError_t launch_tlwe_lincomb<uint32_t>(const uint32_t *, const uint32_t *, int, int,
                                                  uint32_t, uint32_t *, int, bool, int, hipStream_t);
template hipError_t launch_tlwe_lincomb<uint64_t>(const uint64_t *, const uint64_t *, int, int,
                                                  uint64_t, uint64_t *, int, bool, int, hipStream_t);

template <class T>
hipError_t launch_tlwe_shl(const T *in, T *out, int W, int shift, int B, hipStream_t s)
{
    if (B <= 0) return hipSuccess;
    const size_t n = size_t(B) * W;
    hipLaunchKernelGGL(k_tlwe_shl<T>, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, in, out,
                       shift, n);
    return hipGetLastError();
}
template hipError_t launch_tlwe_shl<uint32_t>(const uint32_t *, uint32_t *, int, int, int,
                                              hipStream_t);
template hipError_t launch_tlwe_shl<uint64_t>(const uint64_t *, uint64_t *, int, int, int,
                                              hipStream_t);

template <class T>
hipError_t launch_tlwe_add_body(T *tlwe, int W, T add, int B, hipStream_t s)
{
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tlwe_add_body<T>, dim3(unsigned((B + 255) / 256)), dim3(256), 0, s, tlwe,
                       W, add, B);
    return hipGetLastError();
}
template hipError_t launch_tlwe_add_body<uint32_t>(uint32_t *, int, uint32_t, int, hipStream_t);
template hipError_t launch_tlwe_add_body<uint64_t>(uint64_t *, int, uint64_t, int, hipStream_t);

template <class T>
hipError_t launch_tv_const(T *tv, T u, int N, hipStream_t s)
{
    hipLaunchKernelGGL(k_tv_const<T>, dim3(unsigned((N + 255) / 256)), dim3(256), 0, s, tv, u, N);
    return hipGetLastError();
}
template hipError_t launch_tv_const<uint32_t>(uint32_t *, uint32_t, int, hipStream_t);
template hipError_t launch_tv_const<uint64_t>(uint64_t *, uint64_t, int, hipStream_t);

template <class T>
hipError_t launch_tv_ide(T *tv, uint32_t scale_bits, int N, hipStream_t s)
{
    hipLaunchKernelGGL(k_tv_ide<T>, dim3(unsigned((N + 255) / 256)), dim3(256), 0, s, tv,
                       scale_bits, N);
    return hipGetLastError();
}
template hipError_t launch_tv_ide<uint32_t>(uint32_t *, uint32_t, int, hipStream_t);
template hipError_t launch_tv_ide<uint64_t>(uint64_t *, uint32_t, int, hipStream_t);

template <int C>
void launch_iks21_tiled(const uint64_t *in, const uint32_t *ksk, uint32_t *out, int tn, int B,
                        hipStream_t s)
{
    constexpr int TJ = KS21::t;
    const dim3 grid(unsigned((B + C - 1) / C), unsigned((tn + 1 + 63) / 64));
    hipLaunchKernelGGL((k_iks_tiled<uint64_t, Lvl2::N, KS21::t, KS21::basebit, C, TJ>), grid,
                       dim3(256), 0, s, in, ksk, out, tn, B);
}

template <int CW>
void launch_iks21_ct(const uint64_t *in, const uint32_t *ksk, uint32_t *out, int B, hipStream_t s)
{
    const dim3 grid(unsigned((B + 63) / 64), unsigned(Lvl1::N / (4 * CW)));
    hipLaunchKernelGGL((k_iks_ct<uint64_t, Lvl2::N, Lvl1::N, KS21::t, KS21::basebit, CW>), grid,
                       dim3(256), 0, s, in, ksk, out, B);
}

hipError_t launch_iks(int kind, const void *in, const uint32_t *ksk, uint32_t *out, int tn, int B,
                      hipStream_t s, int tile)
{
    if (B <= 0) return hipSuccess;
    switch (kind) {
    case 10:
    case 20: {
        // `tile` = workgroups per ciphertext over the digit positions
        const int split = tile == 2 || tile == 4 || tile == 8 ? tile : 1;
        if (kind == 10)
            launch_iks_rows<uint32_t, Lvl1::N, KS10::t, KS10::basebit>(
                static_cast<const uint32_t *>(in), ksk, out, tn, B, split, s);
        else
            launch_iks_rows<uint64_t, Lvl2::N, KS20::t, KS20::basebit>(
                static_cast<const uint64_t *>(in), ksk, out, tn, B, split, s);
        break;
    }
    case 21: {
        // `tile` ciphertexts per workgroup: 8, 16 or 32 (0 = 32) in the
        // column-per-lane form, all 10 digit rows of a mask word in flight
        // (larger tiles read each key row for more ciphertexts, smaller ones
        // give small batches more workgroups); 64 in the ciphertext-per-lane
        // form (LDS-staged rows, 16-byte reads; 4-5x faster from 1024
        // ciphertexts up, slower below 512, where a workgroup's 64 lanes are
        // mostly idle).  The second form is built for the lvl1 output width.
        const auto *x = static_cast<const uint64_t *>(in);
        // the second form with 16 output columns per wave, or 32 above 4096
        // ciphertexts (measured equal at 4096, 2 % faster at 8192 and 16 %
        // at 16384, where the wider slice's halved key traffic pays for its
        // third wave per SIMD instead of a fourth)
        if (tile == 64 && tn == Lvl1::N && B > 4096)
            launch_iks21_ct<32>(x, ksk, out, B, s);
        else if (tile == 64 && tn == Lvl1::N)
            launch_iks21_ct<16>(x, ksk, out, B, s);
        else if (tile == 8)
            launch_iks21_tiled<8>(x, ksk, out, tn, B, s);
        else if (tile == 16)
            launch_iks21_tiled<16>(x, ksk, out, tn, B, s);
        else
            launch_iks21_tiled<32>(x, ksk, out, tn, B, s);
        break;
    }
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_tlwe1_lincomb(const uint32_t *a, const uint32_t *b, int sa, int sb,
                                uint32_t offset, uint32_t *out, int B, hipStream_t s)
{
    if (B <= 0) return hipSuccess;
    const size_t n = size_t(B) * (Lvl1::N + 1);
    hipLaunchKernelGGL(k_tlwe1_lincomb, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, a, b, sa,
                       sb, offset, out, B);
    return hipGetLastError();
}

hipError_t launch_tlwe1_add_body(uint32_t *tlwe, uint32_t add, int B, hipStream_t s)
{
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tlwe1_add_body, dim3(unsigned((B + 255) / 256)), dim3(256), 0, s, tlwe,
                       add, B);
    return hipGetLastError();
}

hipError_t launch_tv1_const(uint32_t *tv, uint32_t u, hipStream_t s)
{
    hipLaunchKernelGGL(k_tv1_const, dim3(Lvl1::N / 256), dim3(256), 0, s, tv, u);
    return hipGetLastError();
}

hipError_t launch_tv1_ide(uint32_t *tv, uint32_t scale_bits, hipStream_t s)
{
    hipLaunchKernelGGL(k_tv1_ide, dim3(Lvl1::N / 256), dim3(256), 0, s, tv, scale_bits);
    return hipGetLastError();
}

}  // namespace fhe
