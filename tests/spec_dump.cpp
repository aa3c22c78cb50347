// Test helper: prints the oracle's tables and the shipped packed spec for every
// (node type code, offset, data type), one line each, for test_tables.py.
#include <cstdio>
#include "../oracle/oracle_tables.h"
#include "../ptilopsis_amd/csrc/rv_spec.h"

int main() {
    static oracle_tables t;
    oracle_tables_init(&t);
    static rv_dense d;
    rv_spec_expand(&d);
    for (int k = 0; k < RV_NTYPES; ++k) {
        int code = RV_TYPE_CODES[k];
        for (int o = 0; o < 4; ++o)
            for (int r = 0; r < 6; ++r) {
                uint32_t b = d.bits[k][o][r];
                printf("%d %d %d  %u %u %u %u %u %u %u %u %u %u  %u %u %u %u %u %u %u %u %u %u\n", code, o, r,
                       t.has_instr[code][o][r] ? 1u : 0u, t.instr[code][o][r], t.get_output[code][o][r],
                       t.operand[code][o][r][0], t.operand[code][o][r][1], t.has_output[code][o][r] ? 1u : 0u,
                       t.parent_arg[code][o][r], t.iconst[code][o], t.jt[code][o], t.size[code][r],
                       b & 1u, d.word[k][o][r], (b >> 4) & 7u, (b >> 7) & 7u, (b >> 10) & 7u, (b >> 1) & 1u,
                       (b >> 2) & 3u, (b >> 13) & 7u, (b >> 16) & 7u, (b >> 19) & 7u);
            }
    }
    for (int p = 0; p < 44; ++p) printf("map %d %u %u\n", p, t.pareas_map[p], RV_FROM_PAREAS[p]);
}
