"""Write the golden vectors under tests/golden/ (run in a container that has
the reference sources, after `make ref`).

Every vector is produced by the reference's own TFHEpp, compiled from its
sources into oracle/_ref/libtfhepp_ref.so, on inputs and keys drawn from this
repository's seeded generators (the reference's own RNG cannot be seeded,
thirdparty/TFHEpp/include/utils.hpp:17-22).  Keys are not stored: the tests
regenerate them from the recorded seeds.  Key-switching keys for the
integer-exact IKS vectors are uniform words defined row by row from a seed
(only the rows a given input selects are ever materialised).

TFHEpp is compiled with params/128bit.hpp, so the lvl0 dimension of every
vector here is n0 = 672.

Usage:  python tests/golden/make_golden.py [many_lut | lvl2]
        (many_lut: write only pbs_many_lvl01_ref672.npz;
         lvl2: write only pbs_lvl02_ref672.npz)
"""

import mmap
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import cufhedb_amd as F  # noqa: E402
import _oracle as O  # noqa: E402

N1, N2 = 1024, 2048


def u32(rng, shape):
    return rng.integers(0, 1 << 32, shape, dtype=np.uint64).astype(np.uint32)


def u64(rng, shape):
    return rng.integers(0, 1 << 64, shape, dtype=np.uint64)


def sparse_ksk(shape, seed, tlwe, W, t, basebit):
    """Uniform-word key-switching key of `shape`, materialising only the rows
    that `tlwe` selects (anonymous mapping: untouched pages stay unbacked)."""
    nbytes = int(np.prod(shape)) * 4
    mm = mmap.mmap(-1, nbytes)
    ksk = np.frombuffer(mm, dtype=np.uint32).reshape(shape)
    dn = shape[0]
    prec = 1 << (W - (1 + basebit * t))
    mask = (1 << basebit) - 1
    for i in range(dn):
        abar = (int(tlwe[i]) + prec) % (1 << W)
        for j in range(t):
            d = (abar >> (W - (j + 1) * basebit)) & mask
            if d:
                row_rng = np.random.default_rng([seed, i, j, d])
                ksk[i, j, d - 1, :] = u32(row_rng, shape[3])
    return ksk, mm


def main():
    ref = O.ref()
    if ref is None:
        sys.exit("oracle/_ref/libtfhepp_ref.so is missing: run `make ref` first")
    assert ref.ref_lvl0_n() == 672
    P = O.P
    out = {}

    # ---- TwistIFFT / TwistFFT / PolyMul (mulfft.hpp:58-140)
    rng = np.random.default_rng(101)
    a1 = u32(rng, (4, N1))
    fd1 = np.zeros((4, N1))
    back1 = np.zeros((4, N1), np.uint32)
    for b in range(4):
        ref.ref_twist_ifft_lvl1(P(fd1[b]), P(a1[b]))
        ref.ref_twist_fft_lvl1(P(back1[b]), P(fd1[b]))
    digits = (rng.integers(0, 64, (4, N1)) - 32).astype(np.int64).astype(np.uint32)
    tor = u32(rng, (4, N1))
    pm = np.zeros((4, N1), np.uint32)
    for b in range(4):
        ref.ref_polymul_lvl1(P(pm[b]), P(np.ascontiguousarray(digits[b])), P(tor[b]))
    np.savez_compressed(os.path.join(HERE, "fft_lvl1.npz"), a=a1, fd=fd1, back=back1,
                        pm_a=digits, pm_b=tor, pm=pm)

    a2 = u64(rng, (2, N2))
    fd2 = np.zeros((2, N2))
    back2 = np.zeros((2, N2), np.uint64)
    for b in range(2):
        ref.ref_twist_ifft_lvl2(P(fd2[b]), P(a2[b]))
        ref.ref_twist_fft_lvl2(P(back2[b]), P(fd2[b]))
    np.savez_compressed(os.path.join(HERE, "fft_lvl2.npz"), a=a2, fd=fd2, back=back2)

    # ---- DecompositionPolynomial (trgsw.cpp:21-39)
    p1 = u32(rng, (2, N1))
    p1[0, :8] = [0, 1, 0xFFFFFFFF, 1 << 31, (1 << 31) - 1, 1 << 13, (1 << 14) - 1, 1 << 14]
    d1 = np.zeros((2, 3, N1), np.uint32)
    for b in range(2):
        for d in range(3):
            ref.ref_decompose_lvl1(P(d1[b, d]), P(p1[b]), d)
    p2 = u64(rng, (2, N2))
    p2[0, :4] = [0, 1, (1 << 64) - 1, 1 << 63]
    d2 = np.zeros((2, 4, N2), np.uint64)
    for b in range(2):
        for d in range(4):
            ref.ref_decompose_lvl2(P(d2[b, d]), P(p2[b]), d)
    np.savez_compressed(os.path.join(HERE, "decomp.npz"), p1=p1, d1=d1, p2=p2, d2=d2)

    # ---- external product (trgsw.cpp:78-103); TRGSW from the seeded keygen
    p672 = F.preset(F.PRESET_REF672)
    sk = F.SecretKeys(p672, 672001)
    ep_in = u32(rng, (3, 2, N1))
    ep_out = np.zeros((3, 3, 2, N1), np.uint32)  # [m index][b]
    ms = np.array([1, 0, -1])
    for mi, m in enumerate(ms):
        trgsw = F.trgsw_encrypt_u32(12, int(m) + 5, sk.s1, int(m), 2.0 ** -25)
        fd = np.zeros((6, 2, N1))
        ref.ref_trgsw_to_fft_lvl1(P(fd), P(trgsw))
        for b in range(3):
            ref.ref_external_product_lvl1(P(ep_out[mi, b]), P(ep_in[b]), P(fd))
    np.savez_compressed(os.path.join(HERE, "extprod_lvl1.npz"), sk_seed=672001, trgsw_seed=12,
                        ms=ms, trlwe=ep_in, out=ep_out)

    ep2_in = u64(rng, (2, 2, N2))
    ep2_out = np.zeros((2, 2, 2, N2), np.uint64)
    ms2 = np.array([1, 0])
    for mi, m in enumerate(ms2):
        trgsw = np.zeros((8, 2, N2), np.uint64)
        F.check(F.lib().fhe_trgsw_encrypt_u64(13, int(m) + 5, F._ptr(sk.s2), int(m), 2.0 ** -52,
                                              F._ptr(trgsw)))
        fd = np.zeros((8, 2, N2))
        ref.ref_trgsw_to_fft_lvl2(P(fd), P(trgsw))
        for b in range(2):
            ref.ref_external_product_lvl2(P(ep2_out[mi, b]), P(ep2_in[b]), P(fd))
    np.savez_compressed(os.path.join(HERE, "extprod_lvl2.npz"), sk_seed=672001, trgsw_seed=13,
                        ms=ms2, trlwe=ep2_in, out=ep2_out)

    # ---- PBS lvl01 (gatebootstrapping.cpp:14-23) on seeded keys
    bk = F.keygen_bk_lvl01(sk, 672002)
    bkfft = np.zeros(bk.shape)
    for i in range(bk.shape[0]):
        ref.ref_trgsw_to_fft_lvl1(P(bkfft[i]), P(np.ascontiguousarray(bk[i])))
    B = 8
    m = rng.integers(0, 8, B)
    lut = rng.integers(0, 8, 8)
    j = np.arange(N1)
    box = (j + 64) >> 7
    tv = (np.where(box == 8, -lut[box % 8], lut[box % 8]).astype(np.int64) << 28)
    tv = (tv % (1 << 32)).astype(np.uint32)
    c0 = F.tlwe_encrypt_u32(672100, 0, sk.s0, (m << 28).astype(np.uint32), p672.alpha0)
    pbs = np.zeros((B, N1 + 1), np.uint32)
    acc = np.zeros((2, 2, N1), np.uint32)
    for b in range(B):
        ref.ref_pbs_lvl01(P(pbs[b]), P(c0[b]), P(bkfft), P(tv))
    for b in range(2):
        ref.ref_blind_rotate_lvl01(P(acc[b]), P(c0[b]), P(bkfft), P(tv))
    np.savez_compressed(os.path.join(HERE, "pbs_lvl01_ref672.npz"), sk_seed=672001, bk_seed=672002,
                        enc_seed=672100, m=m, lut=lut, tv=tv, c0=c0, out=pbs, acc=acc,
                        out_phase=F.tlwe_phase_u32(pbs, sk.s1))

    # ---- IdentityKeySwitch lvl10 / lvl20 / lvl21 on uniform-word keys (integer exact)
    iks = {}
    tl1 = u32(rng, (2, N1 + 1))
    o10 = np.zeros((2, 673), np.uint32)
    for b in range(2):
        ksk, mm = sparse_ksk((N1, 2, 1023, 673), 1010 + b, tl1[b], 32, 2, 10)
        ref.ref_iks_lvl10(P(o10[b]), P(tl1[b]), P(ksk))
        del ksk
        mm.close()
    iks.update(in10=tl1, out10=o10)
    tl2 = u64(rng, (2, N2 + 1))
    o20 = np.zeros((2, 673), np.uint32)
    o21 = np.zeros((2, N1 + 1), np.uint32)
    for b in range(2):
        ksk, mm = sparse_ksk((N2, 2, 1023, 673), 2020 + b, tl2[b], 64, 2, 10)
        ref.ref_iks_lvl20(P(o20[b]), P(tl2[b]), P(ksk))
        del ksk
        mm.close()
        ksk, mm = sparse_ksk((N2, 10, 7, N1 + 1), 2121 + b, tl2[b], 64, 10, 3)
        ref.ref_iks_lvl21(P(o21[b]), P(tl2[b]), P(ksk))
        del ksk
        mm.close()
    iks.update(in2=tl2, out20=o20, out21=o21)
    np.savez_compressed(os.path.join(HERE, "iks.npz"), seeds10=np.array([1010, 1011]),
                        seeds20=np.array([2020, 2021]), seeds21=np.array([2121, 2122]), **iks)

    # ---- gates on real keys: HomNAND (gate.cpp:60-65), HEDB MSB / Ide bootstrapping
    ksk10 = F.keygen_ksk(sk, 10, 672003)
    G = 16
    bits = rng.integers(0, 2, (G, 2))
    mu = 1 << 29
    ca = F.tlwe_encrypt_u32(672200, 0, sk.s1, np.where(bits[:, 0] == 1, mu, -mu % (1 << 32)),
                            p672.alpha1)
    cb = F.tlwe_encrypt_u32(672201, 0, sk.s1, np.where(bits[:, 1] == 1, mu, -mu % (1 << 32)),
                            p672.alpha1)
    nand = np.zeros((G, N1 + 1), np.uint32)
    for g in range(G):
        ref.ref_hom_nand_lvl1(P(nand[g]), P(ca[g]), P(cb[g]), P(bkfft), P(ksk10))
    # MSB bootstrapping of 4-bit values filling the torus (v * 2^28), identity
    # bootstrapping of 4-bit values under a padding bit (v * 2^27)
    vals = rng.integers(0, 16, G)
    cm = F.tlwe_encrypt_u32(672203, 0, sk.s1, (vals << 28).astype(np.uint32), p672.alpha1)
    cv = F.tlwe_encrypt_u32(672202, 0, sk.s1, (vals << 27).astype(np.uint32), p672.alpha1)
    msb_logic = np.zeros((G, N1 + 1), np.uint32)
    msb_arith = np.zeros((G, N1 + 1), np.uint32)
    ide = np.zeros((G, N1 + 1), np.uint32)
    for g in range(G):
        ref.ref_msb_gate_lvl1(P(msb_logic[g]), P(cm[g]), P(bkfft), P(ksk10), 1)
        ref.ref_msb_gate_lvl1(P(msb_arith[g]), P(cm[g]), P(bkfft), P(ksk10), 0)
        ref.ref_ide_gate_lvl1(P(ide[g]), P(cv[g]), F.ctypes.c_uint32(27), P(bkfft), P(ksk10))
    np.savez_compressed(os.path.join(HERE, "gates_ref672.npz"), sk_seed=672001, bk_seed=672002,
                        ksk_seed=672003, bits=bits, ca=ca, cb=cb, nand=nand,
                        nand_phase=F.tlwe_phase_u32(nand, sk.s1), vals=vals, cv=cv, cm=cm,
                        msb_logic=F.tlwe_phase_u32(msb_logic, sk.s1),
                        msb_arith=F.tlwe_phase_u32(msb_arith, sk.s1),
                        ide=F.tlwe_phase_u32(ide, sk.s1))
    many_lut(ref, sk, p672, bkfft)
    lvl2_pbs(ref, sk, p672)
    print("golden vectors written to", HERE)


def many_lut(ref, sk, p672, bkfft):
    """GateBootstrappingManyLUT<lvl01param, K> (gatebootstrapping.hpp:119-130)
    for K = 2 (Z_8 messages) and K = 4 (Z_4 messages: the stride-4 modulus
    switch quadruples the rounding noise), random tables."""
    P = O.P
    rng = np.random.default_rng(777)
    vec = {"sk_seed": 672001, "bk_seed": 672002}
    for K, bits in ((2, 3), (4, 2)):
        B = 6
        m = rng.integers(0, 1 << bits, B)
        luts = rng.integers(0, 16, (K, 1 << bits))
        tv = F.many_lut_testvector(luts << 28, bits, N1)
        c0 = F.tlwe_encrypt_u32(672300 + K, 0, sk.s0, (m << (31 - bits)).astype(np.uint32),
                                p672.alpha0)
        out = np.zeros((B, K, N1 + 1), np.uint32)
        for b in range(B):
            assert ref.ref_pbs_many_lvl01(P(out[b]), P(c0[b]), P(bkfft), P(tv), K) == 0
        vec.update({f"m{K}": m, f"luts{K}": luts, f"tv{K}": tv, f"c0_{K}": c0, f"out{K}": out,
                    f"phase{K}": F.tlwe_phase_u32(out.reshape(-1, N1 + 1), sk.s1).reshape(B, K)})
    np.savez_compressed(os.path.join(HERE, "pbs_many_lvl01_ref672.npz"), **vec)


def z8_testvector64(lut, N=N2):
    """Z_8 lookup table on the 64-bit torus: N/8 = 256 entries per value,
    half a box of offset (the lvl02 form of the lvl01 tables above)."""
    j = np.arange(N)
    box = (j + N // 16) // (N // 8)
    vals = np.array(lut, np.uint64)[box % 8] << np.uint64(60)
    return np.where(box == 8, np.uint64(0) - vals, vals).astype(np.uint64)


def lvl2_pbs(ref, sk, p672):
    """GateBootstrappingTLWE2TLWEFFT<lvl02param> (gatebootstrapping.cpp:14-23)
    and GateBootstrappingManyLUT<lvl02param, 2> (gatebootstrapping.hpp:119-130)
    on a seeded lvl02 key (BK seed 672003), Z_8 plaintexts m * 2^28, random
    tables; phases with the lvl2 key."""
    P = O.P
    bk2 = F.keygen_bk_lvl02(sk, 672003)
    bkfft = np.zeros(bk2.shape)
    for i in range(bk2.shape[0]):
        ref.ref_trgsw_to_fft_lvl2(P(bkfft[i]), P(np.ascontiguousarray(bk2[i])))
    rng = np.random.default_rng(2002)
    B = 8
    m = rng.integers(0, 8, B)
    lut = rng.integers(0, 8, 8)
    tv = z8_testvector64(lut)
    c0 = F.tlwe_encrypt_u32(672400, 0, sk.s0, (m << 28).astype(np.uint32), p672.alpha0)
    out = np.zeros((B, N2 + 1), np.uint64)
    for b in range(B):
        ref.ref_pbs_lvl02(P(out[b]), P(c0[b]), P(bkfft), P(tv))
    # two tables, Z_4 plaintexts m * 2^29 (the stride-2 modulus switch)
    mm = rng.integers(0, 4, 4)
    luts = rng.integers(0, 16, (2, 4))
    tvm = F.many_lut_testvector(luts.astype(np.uint64) << np.uint64(60), 2, N2)
    cm = F.tlwe_encrypt_u32(672401, 0, sk.s0, (mm << 29).astype(np.uint32), p672.alpha0)
    outm = np.zeros((4, 2, N2 + 1), np.uint64)
    for b in range(4):
        assert ref.ref_pbs_many_lvl02(P(outm[b]), P(cm[b]), P(bkfft), P(tvm), 2) == 0
    np.savez_compressed(os.path.join(HERE, "pbs_lvl02_ref672.npz"), sk_seed=672001,
                        bk_seed=672003, m=m, lut=lut, tv=tv, c0=c0,
                        out_phase=O.tlwe_phase64(out, sk.s2), mm=mm, luts=luts, tvm=tvm, cm=cm,
                        outm_phase=O.tlwe_phase64(outm.reshape(-1, N2 + 1), sk.s2).reshape(4, 2))


def lvl2_only():
    """Write only the lvl02 vectors (the other files are unchanged)."""
    ref = O.ref()
    if ref is None:
        sys.exit("oracle/_ref/libtfhepp_ref.so is missing: run `make ref` first")
    p672 = F.preset(F.PRESET_REF672)
    lvl2_pbs(ref, F.SecretKeys(p672, 672001), p672)


def many_lut_only():
    """Write only the many-table vectors (the other files are unchanged)."""
    ref = O.ref()
    if ref is None:
        sys.exit("oracle/_ref/libtfhepp_ref.so is missing: run `make ref` first")
    p672 = F.preset(F.PRESET_REF672)
    sk = F.SecretKeys(p672, 672001)
    bk = F.keygen_bk_lvl01(sk, 672002)
    bkfft = np.zeros(bk.shape)
    for i in range(bk.shape[0]):
        ref.ref_trgsw_to_fft_lvl1(O.P(bkfft[i]), O.P(np.ascontiguousarray(bk[i])))
    many_lut(ref, sk, p672, bkfft)


if __name__ == "__main__":
    if sys.argv[1:] == ["many_lut"]:
        many_lut_only()
    elif sys.argv[1:] == ["lvl2"]:
        lvl2_only()
    else:
        main()
