/*
 * libgmpt.h -- the go-ethereum C-ABI, same symbols and signatures as the
 * reference's mpt-with-compress/lib/libgmpt.h:1-55 (implemented by
 * libgmpt_amd.so on top of gmpt_amd.h).
 *
 * Differences from the reference, each deliberate:
 *  - build_mpt_2phase runs PhaseNU as its name says; the reference runs LockNU
 *    there (libgmpt.cu:107-111).  Both produce the same trie and root.
 *  - insert_mpt_2phase is implemented (an empty stub in libgmpt.cu:240-246).
 *  - The device ordinal comes from the GMPT_DEVICE environment variable
 *    (default 0) instead of a hard-coded device 1 (libgmpt.cu:27).
 *  - Nothing is printed; errors return NULL (or 0 from preprocess) and
 *    gmpt_last_error() (gmpt_amd.h) tells why, instead of exit().
 *  - get_all_nodes keeps the transaction and receipt tries alive.
 *  - Returned root hashes are 32-byte new[] buffers owned by the caller, as in
 *    the reference; get_all_nodes arrays too.
 */
#ifndef LIBGMPT_H
#define LIBGMPT_H

#ifdef __cplusplus
extern "C" {
#endif

#include "stdint.h"

enum TrieType {
  STATE_TRIE = 0,
  TRANSACTION_TRIE = 1,
  RECEIPT_TRIE = 2,
};

struct nodeset {
  const uint8_t *hashs;
  const uint8_t *encs;
  const unsigned long long *encs_indexs;
  unsigned long long num;
};

/* libgmpt.h:21-24 / libgmpt.cu:45-129 */
const uint8_t *build_mpt_2phase(enum TrieType trie_type, const uint8_t *keys_hexs, int *keys_hexs_indexs,
                                const uint8_t *values_bytes, int64_t *values_bytes_indexs,
                                const uint8_t **values_hps, int insert_num);

/* libgmpt.h:26-29 / libgmpt.cu:131-214 */
const uint8_t *build_mpt_olc(enum TrieType trie_type, const uint8_t *keys_hexs, int *keys_hexs_indexs,
                             const uint8_t *values_bytes, int64_t *values_bytes_indexs,
                             const uint8_t **values_hps, int insert_num);

/* libgmpt.h:31 / libgmpt.cu:25-43 */
int preprocess();

/* libgmpt.h:33-34 / libgmpt.cu:216-238 */
struct nodeset get_all_nodes(enum TrieType trie_type, const uint8_t *keys_hexs, int *keys_hexs_indexs, int num);

/* libgmpt.h:38 / libgmpt.cu:323-327 */
void *init_mpt();

/* libgmpt.h:40-44 / libgmpt.cu:240-246 */
const uint8_t *insert_mpt_2phase(void *mpt, const uint8_t *keys_hexs, int *keys_hexs_indexs,
                                 const uint8_t *values_bytes, int64_t *values_bytes_indexs,
                                 const uint8_t **values_hps, int insert_num);
/* libgmpt.h:45-49 / libgmpt.cu:248-321 */
const uint8_t *insert_mpt_olc(void *mpt, const uint8_t *keys_hexs, int *keys_hexs_indexs,
                              const uint8_t *values_bytes, int64_t *values_bytes_indexs,
                              const uint8_t **values_hps, int insert_num);

/* libgmpt.h:51 / libgmpt.cu:329-331 */
void free_mpt(void *mpt);

#ifdef __cplusplus
}
#endif

#endif /* LIBGMPT_H */
