/* Poseidon hash chains and their circuit witness (CPU oracle; test
 * infrastructure only).
 *
 * Restates the reference's width-3 Poseidon permutation with 8 full and 60
 * partial rounds and the x^5 S-box (bls12_381_g1_msm.cu:2089-2157), the
 * dependent hash chain built on it (:2159-2216), and the scatter of the
 * chain's witness stream into the eight-column advice layout (:2018-2072).
 * MDS (row-major 3x3) and round constants (68 x 3) come from the caller.
 *
 * Pinned in tests/test_oracle_poseidon.py by a textbook HADES restatement in
 * Python integers (add-round-key, S-box, MDS per round), by the inverse
 * permutation (round trip), and by building the advice columns directly
 * from the permutation's rows (poseidon_cpu.rs:197-263 order) against the
 * scatter here. */
#include <string.h>

#include "oracle.h"

#define PW 3
#define PROUNDS 68

static int is_full(unsigned round) { return round < 4 || round >= 64; }

/* One permutation of `state`; when `w` is not NULL it receives the 128
 * witness values in the circuit's order: per full round the three cubes then
 * the three post-MDS values; per batch of six partial rounds the last lane of
 * the first five and then the whole state after the sixth. */
void orc_poseidon_permutation(orc_fr state[3], const orc_fr *mds, const orc_fr *rc, orc_fr *w)
{
    size_t off = 0;
    for (int l = 0; l < PW; ++l) orc_fr_add(&state[l], &state[l], &rc[l]);
    for (unsigned round = 0; round < PROUNDS; ++round) {
        orc_fr s[PW], next[PW];
        for (int l = 0; l < PW; ++l) {
            s[l] = state[l];
            if (is_full(round) || l == 2) {
                orc_fr sq, cube;
                orc_fr_mul(&sq, &state[l], &state[l]);
                orc_fr_mul(&cube, &sq, &state[l]);
                if (w && is_full(round)) w[off + l] = cube;
                orc_fr_mul(&s[l], &cube, &sq);
            }
        }
        for (int l = 0; l < PW; ++l) {
            if (round + 1 < PROUNDS) next[l] = rc[(round + 1) * PW + l];
            else memset(&next[l], 0, sizeof next[l]);
            for (int c = 0; c < PW; ++c) {
                orc_fr t;
                orc_fr_mul(&t, &mds[l * PW + c], &s[c]);
                orc_fr_add(&next[l], &next[l], &t);
            }
        }
        if (w) {
            if (is_full(round)) {
                for (int l = 0; l < PW; ++l) w[off + 3 + l] = next[l];
                off += 6;
            } else {
                unsigned in_batch = (round - 4) % 6;
                if (in_batch < 5) {
                    w[off + in_batch] = next[2];
                } else {
                    for (int l = 0; l < PW; ++l) w[off + 5 + l] = next[l];
                    off += 8;
                }
            }
        }
        for (int l = 0; l < PW; ++l) state[l] = next[l];
    }
}

void orc_poseidon_hash_chain(orc_fr *out, orc_fr *witness, size_t len, const orc_fr *initial,
                             const orc_fr *tail0, const orc_fr *tail1, const orc_fr *input_len,
                             const orc_fr *mds, const orc_fr *rc)
{
    orc_fr previous = *initial;
    for (size_t i = 0; i < len; ++i) {
        orc_fr st[PW];
        memset(st, 0, sizeof st);
        st[2] = *input_len;
        orc_fr_add(&st[0], &st[0], &previous);
        orc_fr_add(&st[1], &st[1], tail0);
        orc_poseidon_permutation(st, mds, rc, witness ? witness + i * 256 : NULL);
        orc_fr_add(&st[0], &st[0], tail1);
        orc_poseidon_permutation(st, mds, rc, witness ? witness + i * 256 + 128 : NULL);
        previous = st[0];
        if (out) out[i] = previous;
    }
}

/* Eight advice columns of domain_len rows each (column-major), starting at
 * zero.  Hash h occupies rows 2 + 41 h ...: its witness goes through the
 * packed layout ((column << 8) | row, relative to the permutation's first
 * row, 2 + 41 h + 21 p for permutation p); the rows holding each
 * permutation's input state and input-plus-first-round-constants, and the
 * row adding tail1, come from metadata = {prefix[3], binding, initial,
 * tail0, tail1, input_len, rc0[3]}.  The witness scatter is written first
 * and the metadata rows over it. */
void orc_poseidon_advice_from_witness(orc_fr *out, const orc_fr *witness, size_t hash_count, size_t domain_len,
                                      const orc_fr *metadata, const uint16_t *layout)
{
    memset(out, 0, 8 * domain_len * sizeof(orc_fr));
    for (size_t idx = 0; idx < hash_count * 256; ++idx) {
        size_t h = idx / 256, local = idx % 256, perm = local / 128;
        uint16_t packed = layout[local % 128];
        size_t col = packed >> 8, row = packed & 0xff;
        out[col * domain_len + 2 + 41 * h + 21 * perm + row] = witness[idx];
    }
    for (size_t h = 0; h < hash_count; ++h) {
        size_t r0 = 2 + 41 * h;
        orc_fr first[3] = {h == 0 ? metadata[4] : witness[(h - 1) * 256 + 253], metadata[5], metadata[7]};
        for (int c = 0; c < 3; ++c) {
            out[c * domain_len + r0] = first[c];
            orc_fr_add(&out[c * domain_len + r0 + 1], &first[c], &metadata[8 + c]);
        }
        const orc_fr *fo = witness + h * 256 + 125;   /* first permutation's output state */
        orc_fr second0;
        orc_fr_add(&second0, &fo[0], &metadata[6]);
        out[r0 + 20] = second0;
        out[domain_len + r0 + 20] = fo[0];
        out[2 * domain_len + r0 + 20] = metadata[6];
        orc_fr second[3] = {second0, fo[1], fo[2]};
        for (int c = 0; c < 3; ++c) {
            out[c * domain_len + r0 + 21] = second[c];
            orc_fr_add(&out[c * domain_len + r0 + 22], &second[c], &metadata[8 + c]);
        }
    }
    if (hash_count) {
        out[0] = metadata[0];
        out[domain_len] = metadata[1];
        out[2 * domain_len] = metadata[2];
        out[1] = metadata[3];
    }
}
