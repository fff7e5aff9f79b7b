/* HTJ2K decode look-up tables for the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY: this file is part of the oracle that the test
 * suite and the CPU-baseline leg of the benchmark use as the checker.  The
 * product library never links it.
 *
 * The context-VLC tables are built from the T.814 codeword rows in
 * ht_vlc_rows.h; the U-VLC tables from the prefix/suffix rules of T.814
 * clause 7.3.6.  Entry layouts follow the decoder in the reference
 * (crates/signinum-j2k-native/src/j2c/ht_block_decode.rs:1085-1201):
 *
 *   VLC entry  : bits 0-2 codeword length, bit 3 u_off, bits 4-7 rho,
 *                bits 8-11 e_1, bits 12-15 e_k; index (c_q << 7) | 7 bits.
 *   UVLC entry : bits 0-2 total prefix length, bits 3-6 total suffix length,
 *                bits 7-9 suffix length of quad 0, bits 10-12 prefix value of
 *                quad 0, bits 13-15 prefix value of quad 1; index
 *                (mode << 6) | 6 bits, mode = u_off0 | u_off1 << 1, plus
 *                mode 4 (table 0 only) when the MEL says both residuals
 *                exceed 2.
 */
#include <stdint.h>
#include <string.h>

#include "ht_vlc_rows.h"
#include "oracle.h"

uint16_t oracle_vlc_table0[1024];
uint16_t oracle_vlc_table1[1024];
uint16_t oracle_uvlc_table0[320];
uint16_t oracle_uvlc_table1[256];

static void build_vlc(uint16_t *table, const unsigned char (*rows)[7], int nrows)
{
    for (int idx = 0; idx < 1024; ++idx) {
        int c_q = idx >> 7, bits = idx & 0x7F;
        uint16_t entry = 0;
        for (int r = 0; r < nrows; ++r) {
            const unsigned char *row = rows[r];
            int len = row[6];
            if (row[0] != c_q || (bits & ((1 << len) - 1)) != row[5])
                continue;
            entry = (uint16_t)(len | (row[2] << 3) | (row[1] << 4) | (row[4] << 8) | (row[3] << 12));
            break;
        }
        table[idx] = entry;
    }
}

/* U-VLC prefix (T.814 Table 3): "1" -> 1, "01" -> 2, "001" -> 3, "000" -> 5,
 * read LSB-first.  Returns the prefix length and stores the prefix value. */
static int uvlc_prefix(int bits, int *value)
{
    if (bits & 1) { *value = 1; return 1; }
    if (bits & 2) { *value = 2; return 2; }
    if (bits & 4) { *value = 3; return 3; }
    *value = 5;
    return 3;
}

static int uvlc_suffix_len(int prefix_value)
{
    return prefix_value < 3 ? 0 : (prefix_value == 3 ? 1 : 5);
}

static uint16_t uvlc_entry(int pre_len, int suf_total, int suf0, int base0, int base1)
{
    return (uint16_t)(pre_len | (suf_total << 3) | (suf0 << 7) | (base0 << 10) | (base1 << 13));
}

static void build_uvlc(uint16_t *table, int initial_row)
{
    int modes = initial_row ? 5 : 4;
    for (int mode = 0; mode < modes; ++mode) {
        for (int bits = 0; bits < 64; ++bits) {
            uint16_t e = 0;
            int v0, v1, l0, l1;
            if (mode == 1) {
                l0 = uvlc_prefix(bits, &v0);
                e = uvlc_entry(l0, uvlc_suffix_len(v0), uvlc_suffix_len(v0), v0, 0);
            } else if (mode == 2) {
                l1 = uvlc_prefix(bits, &v1);
                e = uvlc_entry(l1, uvlc_suffix_len(v1), 0, 0, v1);
            } else if (mode == 3) {
                l0 = uvlc_prefix(bits, &v0);
                if (initial_row && v0 > 2) {
                    /* first row, both residuals present, MEL event 0: a
                     * prefix above 2 for quad 0 leaves one bit for quad 1,
                     * whose residual is then 1 + bit. */
                    int b = (bits >> l0) & 1;
                    e = uvlc_entry(l0 + 1, uvlc_suffix_len(v0), uvlc_suffix_len(v0), v0, b + 1);
                } else {
                    l1 = uvlc_prefix(bits >> l0, &v1);
                    e = uvlc_entry(l0 + l1, uvlc_suffix_len(v0) + uvlc_suffix_len(v1),
                                   uvlc_suffix_len(v0), v0, v1);
                }
            } else if (mode == 4) {
                /* first row, MEL event 1: both residuals are coded minus 2. */
                l0 = uvlc_prefix(bits, &v0);
                l1 = uvlc_prefix(bits >> l0, &v1);
                e = uvlc_entry(l0 + l1, uvlc_suffix_len(v0) + uvlc_suffix_len(v1),
                               uvlc_suffix_len(v0), v0 + 2, v1 + 2);
            }
            table[mode * 64 + bits] = e;
        }
    }
}

static int tables_ready = 0;

void oracle_build_tables(void)
{
    if (tables_ready)
        return;
    build_vlc(oracle_vlc_table0, ORACLE_VLC_ROWS0,
              (int)(sizeof(ORACLE_VLC_ROWS0) / sizeof(ORACLE_VLC_ROWS0[0])));
    build_vlc(oracle_vlc_table1, ORACLE_VLC_ROWS1,
              (int)(sizeof(ORACLE_VLC_ROWS1) / sizeof(ORACLE_VLC_ROWS1[0])));
    build_uvlc(oracle_uvlc_table0, 1);
    build_uvlc(oracle_uvlc_table1, 0);
    tables_ready = 1;
}

void oracle_get_tables(uint16_t *vlc0, uint16_t *vlc1, uint16_t *uvlc0, uint16_t *uvlc1)
{
    oracle_build_tables();
    memcpy(vlc0, oracle_vlc_table0, sizeof(oracle_vlc_table0));
    memcpy(vlc1, oracle_vlc_table1, sizeof(oracle_vlc_table1));
    memcpy(uvlc0, oracle_uvlc_table0, sizeof(oracle_uvlc_table0));
    memcpy(uvlc1, oracle_uvlc_table1, sizeof(oracle_uvlc_table1));
}
