/* Internal declarations shared by the oracle's translation units.
 * TEST INFRASTRUCTURE ONLY (see oracle.h). */
#ifndef SGN_ORACLE_INTERNAL_H
#define SGN_ORACLE_INTERNAL_H
#include <stdint.h>
#include <stddef.h>
extern uint16_t oracle_vlc_table0[1024];
extern uint16_t oracle_vlc_table1[1024];
extern uint16_t oracle_uvlc_table0[320];
extern uint16_t oracle_uvlc_table1[256];
#endif
