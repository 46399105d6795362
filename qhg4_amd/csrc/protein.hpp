// AssortativePair's match value on the device: an agent's proteome as the
// reference's ProteinBuilder::translateGenomeHash builds it
// (QHG4/genes/ProteinBuilder.cpp:163-220, getNextNuc / getNextAA :231-289),
// one MD5 digest per protein, and ProteomeComparator::countProtHashMatches
// (QHG4/genes/ProteomeComparator.cpp:25-60): the number of equal digest
// pairs of two agents.
//
// Translation: each strand is read in codons from its first nucleotide
// (nucleotide i of a genome word in bits 2i, 2i + 1; codon index
// n0 * 16 + n1 * 4 + n2).  Outside a protein, a Met codon starts one; inside,
// a stop codon or the strand's end closes it (a partial codon at the end is
// dropped).  A protein's bytes are the amino-acid numbers after its start
// codon, up to the closing codon.  Strand 1 is read after strand 0, from
// its own first codon, and no protein spans the two.
//
// MD5 follows RFC 1321 (little-endian words, padding with 0x80 and the bit
// length); the digest's 16 bytes are the state words A, B, C, D in order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qhg {

constexpr int kProtBudget = 16;     // digests kept per agent slot (more fails the step)
constexpr int32_t kErrProteins = 16;   // an agent has more proteins than kProtBudget
constexpr uint8_t kAaMet = 10, kAaStp = 20;

// the genetic code in the reference's numbering (aTranslationTable,
// ProteinBuilder.h:64-78): codon index -> amino acid 0..19, 20 = stop
__constant__ uint8_t kAminoOf[64] = {
    8, 11, 8, 11, 16, 16, 16, 16, 14, 15, 14, 15, 7, 7, 10, 7,
    13, 6, 13, 6, 12, 12, 12, 12, 14, 14, 14, 14, 9, 9, 9, 9,
    3, 2, 3, 2, 0, 0, 0, 0, 5, 5, 5, 5, 17, 17, 17, 17,
    20, 19, 20, 19, 15, 15, 15, 15, 20, 1, 18, 1, 9, 4, 9, 4,
};

// RFC 1321: T[i] = floor(2^32 |sin(i + 1)|)
__constant__ uint32_t kMd5T[64] = {
    0xd76aa478u, 0xe8c7b756u, 0x242070dbu, 0xc1bdceeeu,
    0xf57c0fafu, 0x4787c62au, 0xa8304613u, 0xfd469501u,
    0x698098d8u, 0x8b44f7afu, 0xffff5bb1u, 0x895cd7beu,
    0x6b901122u, 0xfd987193u, 0xa679438eu, 0x49b40821u,
    0xf61e2562u, 0xc040b340u, 0x265e5a51u, 0xe9b6c7aau,
    0xd62f105du, 0x02441453u, 0xd8a1e681u, 0xe7d3fbc8u,
    0x21e1cde6u, 0xc33707d6u, 0xf4d50d87u, 0x455a14edu,
    0xa9e3e905u, 0xfcefa3f8u, 0x676f02d9u, 0x8d2a4c8au,
    0xfffa3942u, 0x8771f681u, 0x6d9d6122u, 0xfde5380cu,
    0xa4beea44u, 0x4bdecfa9u, 0xf6bb4b60u, 0xbebfbc70u,
    0x289b7ec6u, 0xeaa127fau, 0xd4ef3085u, 0x04881d05u,
    0xd9d4d039u, 0xe6db99e5u, 0x1fa27cf8u, 0xc4ac5665u,
    0xf4292244u, 0x432aff97u, 0xab9423a7u, 0xfc93a039u,
    0x655b59c3u, 0x8f0ccc92u, 0xffeff47du, 0x85845dd1u,
    0x6fa87e4fu, 0xfe2ce6e0u, 0xa3014314u, 0x4e0811a1u,
    0xf7537e82u, 0xbd3af235u, 0x2ad7d2bbu, 0xeb86d391u,
};

__device__ __forceinline__ uint32_t rotl32(uint32_t v, int s) { return (v << s) | (v >> (32 - s)); }

// one MD5 compression of the 16-word block m into st[4]
__device__ __forceinline__ void md5_block(uint32_t (&st)[4], const uint32_t (&m)[16]) {
    constexpr int kS[4][4] = {{7, 12, 17, 22}, {5, 9, 14, 20}, {4, 11, 16, 23}, {6, 10, 15, 21}};
    uint32_t a = st[0], b = st[1], c = st[2], d = st[3];
#pragma unroll
    for (int i = 0; i < 64; i++) {
        uint32_t f;
        int g;
        if (i < 16) { f = (b & c) | (~b & d); g = i; }
        else if (i < 32) { f = (d & b) | (~d & c); g = (5 * i + 1) & 15; }
        else if (i < 48) { f = b ^ c ^ d; g = (3 * i + 5) & 15; }
        else { f = c ^ (b | ~d); g = (7 * i) & 15; }
        const uint32_t t = d;
        d = c;
        c = b;
        b = b + rotl32(a + f + kMd5T[i] + m[g], kS[i >> 4][i & 3]);
        a = t;
    }
    st[0] += a;
    st[1] += b;
    st[2] += c;
    st[3] += d;
}

// amino acid of codon k of a strand (words w)
__device__ __forceinline__ uint32_t codon_amino(const uint64_t *w, int32_t k) {
    uint32_t idx = 0;
#pragma unroll
    for (int t = 0; t < 3; t++) {
        const int32_t pos = 3 * k + t;
        idx = (idx << 2) | (uint32_t)((w[pos >> 5] >> (2 * (pos & 31))) & 3u);
    }
    return kAminoOf[idx];
}

// MD5 of the protein of codons [s, e) of a strand
__device__ __forceinline__ void protein_md5(const uint64_t *w, int32_t s, int32_t e, uint32_t (&st)[4]) {
    st[0] = 0x67452301u;
    st[1] = 0xefcdab89u;
    st[2] = 0x98badcfeu;
    st[3] = 0x10325476u;
    const int32_t len = e - s;
    const int32_t nblk = (len + 8) / 64 + 1;
    for (int32_t b = 0; b < nblk; b++) {
        uint32_t m[16];
#pragma unroll
        for (int q = 0; q < 16; q++) {
            uint32_t word = 0;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int32_t j = 64 * b + 4 * q + t;
                uint32_t by = 0;
                if (j < len) by = codon_amino(w, s + j);
                else if (j == len) by = 0x80u;
                word |= by << (8 * t);
            }
            m[q] = word;
        }
        if (b == nblk - 1) {
            const uint64_t bits = (uint64_t)len * 8u;
            m[14] = (uint32_t)bits;
            m[15] = (uint32_t)(bits >> 32);
        }
        md5_block(st, m);
    }
}

// Digests of an agent's proteins (genome g, gwords words, genome_size
// nucleotides per strand) into out[4 * k] (at most kProtBudget); returns
// the number of proteins, which may exceed the budget (then only the first
// kProtBudget are written).  A genome of 3 nucleotides or fewer has none
// (ProteinBuilder::init refuses it).
__device__ __forceinline__ int32_t protein_digests(const uint64_t *g, int32_t gwords, int32_t genome_size,
                                                   uint32_t *out) {
    if (genome_size <= 3) return 0;
    const int32_t nb = gwords / 2;
    const int32_t C = genome_size / 3;
    int32_t np = 0;
    for (int a = 0; a < 2; a++) {
        const uint64_t *w = g + (int64_t)a * nb;
        bool building = false;
        int32_t start = 0;
        for (int32_t k = 0; k <= C; k++) {
            // codon C stands for the strand's end
            const uint32_t aa = (k < C) ? codon_amino(w, k) : kAaStp;
            if (!building) {
                if (k < C && aa == kAaMet) {
                    building = true;
                    start = k + 1;
                }
            } else if (aa == kAaStp) {
                if (np < kProtBudget) {
                    uint32_t st[4];
                    protein_md5(w, start, k, st);
#pragma unroll
                    for (int q = 0; q < 4; q++) out[4 * np + q] = st[q];
                }
                np++;
                building = false;
            }
        }
    }
    return np;
}

// countProtHashMatches: equal digest pairs of two agents' lists
__device__ __forceinline__ int32_t protein_matches(const uint32_t *a, int32_t na, const uint32_t *b, int32_t nb) {
    int32_t n = 0;
    for (int32_t i = 0; i < na; i++) {
        const uint32_t a0 = a[4 * i], a1 = a[4 * i + 1], a2 = a[4 * i + 2], a3 = a[4 * i + 3];
        for (int32_t j = 0; j < nb; j++)
            n += (a0 == b[4 * j] && a1 == b[4 * j + 1] && a2 == b[4 * j + 2] && a3 == b[4 * j + 3]) ? 1 : 0;
    }
    return n;
}

}  // namespace qhg
