// Host build of the prover's Poseidon2 (whirlaway_amd/csrc/kb.h: the templates
// the kernels instantiate, minus the lane-pair variants) for CPU tests.
// stdin: "perm W n", then 4W + 4W + RP round constants and n states, all
// canonical -> n permuted states, canonical.
// stdin: "mix W n", then n states as raw words -> n outputs of one internal
// linear layer (internal_layer_mix) as raw words, lazy ranges left as they are.
#include <cstdio>
#include <cstring>
#include "../../whirlaway_amd/csrc/kb.h"

static bool read_u32(uint32_t &v) { return scanf("%u", &v) == 1; }

template <int W, int RP, class C>
static int perm(int n) {
    C k;
    uint32_t v;
    for (int r = 0; r < 4; r++) for (int i = 0; i < W; i++) { if (!read_u32(v)) return 2; k.ext_init[r][i] = kb::to_mont(v % kb::P); }
    for (int r = 0; r < 4; r++) for (int i = 0; i < W; i++) { if (!read_u32(v)) return 2; k.ext_term[r][i] = kb::to_mont(v % kb::P); }
    for (int r = 0; r < RP; r++) { if (!read_u32(v)) return 2; k.internal[r] = kb::to_mont(v % kb::P); }
    for (int t = 0; t < n; t++) {
        kb::F s[W];
        for (int i = 0; i < W; i++) { if (!read_u32(v) || v >= kb::P) return 2; s[i] = kb::to_mont(v); }
        kb::p2_permute<W, RP>(s, k, nullptr);
        for (int i = 0; i < W; i++) printf("%u%c", kb::from_mont(s[i]), i + 1 < W ? ' ' : '\n');
    }
    return 0;
}

template <int W>
static int mix(int n) {
    for (int t = 0; t < n; t++) {
        kb::F s[W];
        for (int i = 0; i < W; i++) if (!read_u32(s[i].v)) return 2;
        kb::internal_layer_mix<W>(s);
        for (int i = 0; i < W; i++) printf("%u%c", s[i].v, i + 1 < W ? ' ' : '\n');
    }
    return 0;
}

int main() {
    char mode[8];
    int width, n;
    if (scanf("%7s %d %d", mode, &width, &n) != 3) return 2;
    if (!strcmp(mode, "perm") && width == 16) return perm<16, 20, kb::P2Consts16>(n);
    if (!strcmp(mode, "perm") && width == 24) return perm<24, 23, kb::P2Consts24>(n);
    if (!strcmp(mode, "mix") && width == 16) return mix<16>(n);
    if (!strcmp(mode, "mix") && width == 24) return mix<24>(n);
    return 2;
}
