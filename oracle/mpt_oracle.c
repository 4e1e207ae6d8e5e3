/*
 * mpt_oracle.c -- CPU oracle for GPU MPT batch insert + root hash.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * It restates, function by function, the reference's plain CPU MPT
 * (CpuMPT::Compress::MPT), which is the reference's own correctness oracle:
 *   insert   : mpt-with-compress/include/mpt/cpu_mpt.cuh:232-378
 *   hash     : mpt-with-compress/include/mpt/cpu_mpt.cuh:380-505
 *   get      : mpt-with-compress/include/mpt/cpu_mpt.cuh:507-579
 *   root     : mpt-with-compress/include/mpt/cpu_mpt.cuh:915-934
 *   encoders : mpt-with-compress/include/mpt/node.cuh:46-112, 226-270
 *   RLP      : mpt-with-compress/include/util/rlp.cuh:9-155
 *   hex/HP   : mpt-with-compress/include/util/utils.cuh:311-396
 * Keccak-256 is the reference's CryptoPP 8.7 Keccak_256 call
 * (mpt-with-compress/include/hash/cpu_hash.h:136-145); CryptoPP is not present
 * here, so Keccak-f[1600] is written from FIPS 202 with the original Keccak
 * pad byte 0x01 (a pad parameter lets tests check it against SHA3-256, which
 * shares the permutation and the 136-byte rate and pads with 0x06).
 *
 * Parity pins (tests/test_oracle.py): Keccak vs hashlib.sha3_256 with pad 0x06
 * over lengths 0..600 and Keccak-256 known answers; RLP and hex-prefix vectors;
 * trie invariants (insertion-order independence, a second independent bulk
 * construction, hand-derived encodings of 1-3 key tries).
 */
#define _POSIX_C_SOURCE 200809L
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* Keccak-f[1600] (FIPS 202 section 3), constants derived from the spec.     */
/* ------------------------------------------------------------------------ */

static uint64_t kRC[24];
static int kRho[25];
static int kConstReady = 0;

/* rc(t) from FIPS 202 Algorithm 5: LFSR x^8 + x^6 + x^5 + x^4 + 1. */
static int keccak_rc_bit(int t) {
  if (t % 255 == 0) return 1;
  unsigned r = 0x80; /* R = 10000000 */
  for (int i = 1; i <= t % 255; i++) {
    unsigned r8 = 0; /* prepend 0: R = 0 || R, 9 bits */
    unsigned R = r; /* bits: R[0] is the MSB of the 8-bit register */
    /* represent R[0..7] as bit (7-k) of r */
    unsigned b[9];
    b[0] = 0;
    for (int k = 0; k < 8; k++) b[k + 1] = (R >> (7 - k)) & 1;
    b[0] ^= b[8];
    b[4] ^= b[8];
    b[5] ^= b[8];
    b[6] ^= b[8];
    for (int k = 0; k < 8; k++) r8 |= b[k] << (7 - k);
    r = r8;
  }
  return (r >> 7) & 1;
}

static void keccak_init_consts(void) {
  if (kConstReady) return;
  /* iota round constants: RC[ir] bit (2^j - 1) = rc(j + 7 ir) */
  for (int ir = 0; ir < 24; ir++) {
    uint64_t rc = 0;
    for (int j = 0; j <= 6; j++)
      if (keccak_rc_bit(j + 7 * ir)) rc |= (uint64_t)1 << ((1 << j) - 1);
    kRC[ir] = rc;
  }
  /* rho offsets (FIPS 202 Algorithm 2): walk (x,y) -> (y, 2x+3y) */
  for (int i = 0; i < 25; i++) kRho[i] = 0;
  int x = 1, y = 0;
  for (int t = 0; t < 24; t++) {
    kRho[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
    int nx = y, ny = (2 * x + 3 * y) % 5;
    x = nx;
    y = ny;
  }
  kConstReady = 1;
}

static inline uint64_t rotl64(uint64_t v, int r) {
  r &= 63;
  return r ? (v << r) | (v >> (64 - r)) : v;
}

/* State lane A[x,y] is st[x + 5y]. */
void orc_keccak_f1600(uint64_t st[25]) {
  keccak_init_consts();
  uint64_t C[5], D[5], B[25];
  for (int ir = 0; ir < 24; ir++) {
    for (int x = 0; x < 5; x++)
      C[x] = st[x] ^ st[x + 5] ^ st[x + 10] ^ st[x + 15] ^ st[x + 20];
    for (int x = 0; x < 5; x++) D[x] = C[(x + 4) % 5] ^ rotl64(C[(x + 1) % 5], 1);
    for (int i = 0; i < 25; i++) st[i] ^= D[i % 5];
    /* rho + pi: B[y, 2x+3y] = rot(A[x,y], r[x,y]) */
    for (int x = 0; x < 5; x++)
      for (int y = 0; y < 5; y++)
        B[y + 5 * ((2 * x + 3 * y) % 5)] = rotl64(st[x + 5 * y], kRho[x + 5 * y]);
    /* chi */
    for (int y = 0; y < 5; y++)
      for (int x = 0; x < 5; x++)
        st[x + 5 * y] = B[x + 5 * y] ^ ((~B[(x + 1) % 5 + 5 * y]) & B[(x + 2) % 5 + 5 * y]);
    /* iota */
    st[0] ^= kRC[ir];
  }
}

/* Sponge with rate 136 bytes, 32-byte output.  pad = 0x01 gives Keccak-256
 * (what the reference computes), pad = 0x06 gives SHA3-256. */
void orc_keccak256_pad(const uint8_t *in, size_t len, uint8_t pad, uint8_t out[32]) {
  uint64_t st[25];
  memset(st, 0, sizeof(st));
  const size_t rate = 136;
  uint8_t block[136];
  while (len >= rate) {
    for (int i = 0; i < 17; i++) {
      uint64_t w = 0;
      for (int b = 0; b < 8; b++) w |= (uint64_t)in[8 * i + b] << (8 * b);
      st[i] ^= w;
    }
    orc_keccak_f1600(st);
    in += rate;
    len -= rate;
  }
  memset(block, 0, rate);
  memcpy(block, in, len);
  block[len] ^= pad;
  block[rate - 1] ^= 0x80;
  for (int i = 0; i < 17; i++) {
    uint64_t w = 0;
    for (int b = 0; b < 8; b++) w |= (uint64_t)block[8 * i + b] << (8 * b);
    st[i] ^= w;
  }
  orc_keccak_f1600(st);
  for (int i = 0; i < 32; i++) out[i] = (uint8_t)(st[i / 8] >> (8 * (i % 8)));
}

void orc_keccak256(const uint8_t *in, size_t len, uint8_t out[32]) {
  orc_keccak256_pad(in, len, 0x01, out);
}

/* Number of Keccak-f permutations the sponge runs for a message of len bytes. */
int orc_keccak_perms(int len) { return len / 136 + 1; }

/* ------------------------------------------------------------------------ */
/* RLP (reference util/rlp.cuh:9-155)                                        */
/* ------------------------------------------------------------------------ */

/* rlp.cuh:9-34 */
static int rlp_putint_size(uint64_t i) {
  for (int size = 1;; size++) {
    i >>= 8;
    if (i == 0) return size;
  }
}

/* rlp.cuh:36-51 */
int orc_rlp_write_bytes_size(const uint8_t *bytes, int n) {
  if (n == 1 && bytes[0] <= 0x7f) return 1;
  if (n < 56) return 1 + n;
  return 1 + rlp_putint_size((uint64_t)n) + n;
}

/* rlp.cuh:53-108: big-endian minimal bytes */
static int rlp_putint(uint8_t *b, uint64_t i) {
  int s = rlp_putint_size(i);
  for (int k = 0; k < s; k++) b[k] = (uint8_t)(i >> (8 * (s - 1 - k)));
  return s;
}

/* rlp.cuh:111-116 */
int orc_rlp_puthead_size(uint64_t size) { return size < 56 ? 1 : 1 + rlp_putint_size(size); }

/* rlp.cuh:118-129 */
int orc_rlp_puthead(uint8_t *buf, uint8_t smalltag, uint8_t largetag, uint64_t size) {
  if (size < 56) {
    buf[0] = (uint8_t)(smalltag + size);
    return 1;
  }
  int ss = rlp_putint(buf + 1, size);
  buf[0] = (uint8_t)(largetag + ss);
  return ss + 1;
}

/* rlp.cuh:132-155 */
int orc_rlp_write_bytes(uint8_t *buf, const uint8_t *bytes, int n) {
  if (n == 1 && bytes[0] <= 0x7f) {
    buf[0] = bytes[0];
    return 1;
  }
  int h;
  if (n < 56) {
    buf[0] = (uint8_t)(0x80 + n);
    h = 1;
  } else {
    int ss = rlp_putint(buf + 1, (uint64_t)n);
    buf[0] = (uint8_t)(0xB7 + ss);
    h = 1 + ss;
  }
  memcpy(buf + h, bytes, (size_t)n);
  return h + n;
}

/* ------------------------------------------------------------------------ */
/* hex / compact (reference util/utils.cuh:311-396)                          */
/* ------------------------------------------------------------------------ */

/* utils.cuh:311-323 (bytes2_size is not consulted, as in the reference) */
static int prefix_len(const uint8_t *a, int na, const uint8_t *b) {
  int m = 0;
  while (m < na) {
    if (a[m] != b[m]) return m;
    m++;
  }
  return m;
}

static int bytes_equal(const uint8_t *a, int na, const uint8_t *b, int nb) {
  if (na != nb) return 0;
  return memcmp(a, b, (size_t)na) == 0;
}

/* utils.cuh:340-349 */
int orc_key_bytes_to_hex(const uint8_t *key, int n, uint8_t *hex) {
  for (int i = 0; i < n; i++) {
    hex[2 * i] = key[i] / 16;
    hex[2 * i + 1] = key[i] % 16;
  }
  hex[2 * n] = 16;
  return 2 * n + 1;
}

/* utils.cuh:361-368 */
int orc_hex_to_compact_size(const uint8_t *hex, int n) {
  if (n > 0 && hex[n - 1] == 16) n -= 1;
  return n / 2 + 1;
}

/* utils.cuh:372-396 (Yellow Paper appendix C hex-prefix) */
int orc_hex_to_compact(const uint8_t *hex, int n, uint8_t *out) {
  uint8_t term = 0;
  if (n > 0 && hex[n - 1] == 16) {
    term = 1;
    n -= 1;
  }
  int sz = n / 2 + 1;
  out[0] = (uint8_t)(term << 5);
  if (n & 1) {
    out[0] |= 1 << 4;
    out[0] |= hex[0];
    hex += 1;
    n -= 1;
  }
  out += 1;
  for (int bi = 0, ni = 0; ni < n; bi += 1, ni += 2) out[bi] = (uint8_t)(hex[ni] << 4 | hex[ni + 1]);
  return sz;
}

/* ------------------------------------------------------------------------ */
/* Trie nodes (reference include/mpt/node.cuh:7-334, CPU half)               */
/* ------------------------------------------------------------------------ */

enum { T_NONE = 0, T_FULL = 1, T_SHORT = 2, T_VALUE = 3 };

typedef struct ONode {
  int type;
  const uint8_t *hash; /* digest or raw encoding (<32) or raw value */
  int hash_size;
} ONode;

typedef struct OFull {
  ONode h;
  ONode *childs[17];
  int dirty;
  uint8_t buffer[32];
} OFull;

typedef struct OShort {
  ONode h;
  const uint8_t *key; /* hex nibbles (terminator 16 included at the end of a leaf key) */
  int key_size;
  ONode *val;
  int dirty;
  uint8_t buffer[32];
  uint8_t *key_compact;
} OShort;

typedef struct OValue {
  ONode h;
  const uint8_t *value;
  int value_size;
  int64_t batch_index; /* position in the caller's batch (for gets) */
} OValue;

/* All allocations of one trie live in an arena list so orc_free can release them. */
typedef struct Chunk {
  struct Chunk *next;
  size_t used, cap;
  uint8_t data[];
} Chunk;

typedef struct Census {
  int64_t n_full, n_short_leaf, n_short_ext, n_value_direct;
  int64_t n_hashed;     /* nodes whose encoding was >= 32 B and were hashed */
  int64_t n_embedded;   /* nodes whose encoding was < 32 B (kept raw) */
  int64_t perms;        /* Keccak-f permutations */
  int64_t enc_bytes;    /* sum of encoding lengths of re-encoded nodes */
} Census;

typedef struct OMPT {
  ONode *root;
  Chunk *chunks;
  Census census; /* hash-work counters since the last orc_census call */
  /* owned copies of every batch's key/value bytes (the reference points into
     the caller's buffers; the oracle copies so callers may free theirs) */
} OMPT;

static void *arena_alloc(OMPT *t, size_t sz) {
  sz = (sz + 15) & ~(size_t)15;
  if (!t->chunks || t->chunks->used + sz > t->chunks->cap) {
    size_t cap = sz > (1u << 22) ? sz : (1u << 22);
    Chunk *c = (Chunk *)malloc(sizeof(Chunk) + cap);
    if (!c) {
      fprintf(stderr, "oracle: out of memory\n");
      abort();
    }
    c->next = t->chunks;
    c->used = 0;
    c->cap = cap;
    t->chunks = c;
  }
  void *p = t->chunks->data + t->chunks->used;
  t->chunks->used += sz;
  memset(p, 0, sz);
  return p;
}

void *orc_new(void) {
  OMPT *t = (OMPT *)calloc(1, sizeof(OMPT));
  keccak_init_consts();
  return t;
}

void orc_free(void *h) {
  OMPT *t = (OMPT *)h;
  if (!t) return;
  Chunk *c = t->chunks;
  while (c) {
    Chunk *n = c->next;
    free(c);
    c = n;
  }
  free(t);
}

/* ------------------------------------------------------------------------ */
/* Insert: dfs_put_baseline (cpu_mpt.cuh:232-346)                             */
/* ------------------------------------------------------------------------ */

typedef struct PutRet {
  ONode *node;
  int dirty;
} PutRet;

static PutRet dfs_put(OMPT *t, ONode *node, const uint8_t *key, int key_size, ONode *value) {
  PutRet r;
  /* cpu_mpt.cuh:238-251: key exhausted */
  if (key_size == 0) {
    if (node != NULL && node->type == T_VALUE) {
      OValue *o = (OValue *)node, *nw = (OValue *)value;
      r.node = value;
      r.dirty = !bytes_equal(o->value, o->value_size, nw->value, nw->value_size);
      return r;
    }
    r.node = value;
    r.dirty = 1;
    return r;
  }
  /* cpu_mpt.cuh:254-263: empty slot -> new short node holding the rest */
  if (node == NULL) {
    OShort *s = (OShort *)arena_alloc(t, sizeof(OShort));
    s->h.type = T_SHORT;
    s->key = key;
    s->key_size = key_size;
    s->val = value;
    s->dirty = 1;
    r.node = &s->h;
    r.dirty = 1;
    return r;
  }
  switch (node->type) {
    case T_SHORT: { /* cpu_mpt.cuh:267-313 */
      OShort *s = (OShort *)node;
      int m = prefix_len(s->key, s->key_size, key);
      if (m == s->key_size) {
        PutRet c = dfs_put(t, s->val, key + m, key_size - m, value);
        s->val = c.node;
        if (c.dirty) s->dirty = 1;
        r.node = node;
        r.dirty = c.dirty;
        return r;
      }
      OFull *b = (OFull *)arena_alloc(t, sizeof(OFull));
      b->h.type = T_FULL;
      b->dirty = 1;
      PutRet o = dfs_put(t, NULL, s->key + (m + 1), s->key_size - (m + 1), s->val);
      b->childs[s->key[m]] = o.node;
      PutRet n = dfs_put(t, NULL, key + (m + 1), key_size - (m + 1), value);
      b->childs[key[m]] = n.node;
      if (m == 0) {
        r.node = &b->h;
        r.dirty = 1;
        return r;
      }
      s->key_size = m;
      s->val = &b->h;
      s->dirty = 1;
      r.node = node;
      r.dirty = 1;
      return r;
    }
    case T_FULL: { /* cpu_mpt.cuh:314-329 */
      OFull *f = (OFull *)node;
      PutRet c = dfs_put(t, f->childs[key[0]], key + 1, key_size - 1, value);
      if (c.dirty) {
        f->childs[key[0]] = c.node;
        f->dirty = 1;
      }
      r.node = node;
      r.dirty = c.dirty;
      return r;
    }
    default:
      fprintf(stderr, "oracle: wrong node type %d on insert\n", node->type);
      abort();
  }
}

/* puts_baseline (cpu_mpt.cuh:348-378).  Keys and values are copied into the
 * trie's arena; returns 0, or -1 on a malformed key (empty). */
int orc_puts(void *h, const uint8_t *keys_hexs, const int *keys_indexs, const uint8_t *values_bytes,
             const int64_t *values_indexs, int n) {
  OMPT *t = (OMPT *)h;
  for (int i = 0; i < n; i++) {
    int ks = keys_indexs[2 * i + 1] - keys_indexs[2 * i] + 1;
    int64_t vs = values_indexs[2 * i + 1] - values_indexs[2 * i] + 1;
    if (ks <= 0 || vs < 0) return -1;
    uint8_t *k = (uint8_t *)arena_alloc(t, (size_t)ks);
    memcpy(k, keys_hexs + keys_indexs[2 * i], (size_t)ks);
    uint8_t *v = (uint8_t *)arena_alloc(t, (size_t)(vs ? vs : 1));
    memcpy(v, values_bytes + values_indexs[2 * i], (size_t)vs);
    OValue *vn = (OValue *)arena_alloc(t, sizeof(OValue));
    vn->h.type = T_VALUE;
    vn->value = v;
    vn->value_size = (int)vs;
    vn->batch_index = i;
    PutRet r = dfs_put(t, t->root, k, ks, &vn->h);
    t->root = r.node;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Encoders (node.cuh:46-112 FullNode, node.cuh:226-270 ShortNode)            */
/* ------------------------------------------------------------------------ */

static int full_encode_size(OFull *f, int *payload) {
  int e = 0;
  for (int i = 0; i < 17; i++) {
    ONode *c = f->childs[i];
    e += c ? orc_rlp_write_bytes_size(c->hash, c->hash_size) : 1;
  }
  *payload = e;
  return e + orc_rlp_puthead_size((uint64_t)e);
}

static int full_encode(OFull *f, uint8_t *buf, int payload) {
  int n = orc_rlp_puthead(buf, 0xC0, 0xF7, (uint64_t)payload);
  for (int i = 0; i < 17; i++) {
    ONode *c = f->childs[i];
    if (c)
      n += orc_rlp_write_bytes(buf + n, c->hash, c->hash_size);
    else
      buf[n++] = 0x80;
  }
  return n;
}

static int short_encode_size(OShort *s, int *payload) {
  int kcs = orc_hex_to_compact_size(s->key, s->key_size);
  int e = orc_rlp_write_bytes_size(s->key_compact, kcs) + orc_rlp_write_bytes_size(s->val->hash, s->val->hash_size);
  *payload = e;
  return e + orc_rlp_puthead_size((uint64_t)e);
}

static int short_encode(OShort *s, uint8_t *buf, int payload) {
  int n = orc_rlp_puthead(buf, 0xC0, 0xF7, (uint64_t)payload);
  int kcs = orc_hex_to_compact_size(s->key, s->key_size);
  n += orc_rlp_write_bytes(buf + n, s->key_compact, kcs);
  n += orc_rlp_write_bytes(buf + n, s->val->hash, s->val->hash_size);
  return n;
}

/* ------------------------------------------------------------------------ */
/* Hash: dfs_hashs_dirty_flag (cpu_mpt.cuh:380-505)                           */
/* ------------------------------------------------------------------------ */


static uint8_t *scratch_buf(size_t n) {
  static uint8_t *buf = NULL;
  static size_t cap = 0;
  if (n > cap) {
    free(buf);
    cap = n * 2 + 1024;
    buf = (uint8_t *)malloc(cap);
  }
  return buf;
}

static void finish_node(Census *cs, ONode *node, uint8_t *buffer, const uint8_t *enc, int enc_size) {
  cs->enc_bytes += enc_size;
  if (enc_size < 32) {
    memcpy(buffer, enc, (size_t)enc_size);
    node->hash_size = enc_size;
    node->hash = buffer;
    cs->n_embedded++;
  } else {
    orc_keccak256(enc, (size_t)enc_size, buffer);
    node->hash_size = 32;
    node->hash = buffer;
    cs->n_hashed++;
    cs->perms += orc_keccak_perms(enc_size);
  }
}

static void dfs_hash(OMPT *t, ONode *node) {
  if (!node) return;
  switch (node->type) {
    case T_SHORT: { /* cpu_mpt.cuh:389-434 */
      OShort *s = (OShort *)node;
      if (!s->dirty) return;
      dfs_hash(t, s->val);
      int kcs = orc_hex_to_compact_size(s->key, s->key_size);
      uint8_t *kc = (uint8_t *)arena_alloc(t, (size_t)kcs);
      orc_hex_to_compact(s->key, s->key_size, kc);
      s->key_compact = kc;
      int payload = 0;
      int es = short_encode_size(s, &payload);
      uint8_t *enc = scratch_buf((size_t)es);
      int w = short_encode(s, enc, payload);
      if (w != es) abort();
      finish_node(&t->census, node, s->buffer, enc, es);
      s->dirty = 0;
      return;
    }
    case T_FULL: { /* cpu_mpt.cuh:436-477 */
      OFull *f = (OFull *)node;
      if (!f->dirty) return;
      for (int i = 0; i < 17; i++) dfs_hash(t, f->childs[i]);
      int payload = 0;
      int es = full_encode_size(f, &payload);
      uint8_t *enc = scratch_buf((size_t)es);
      int w = full_encode(f, enc, payload);
      if (w != es) abort();
      finish_node(&t->census, node, f->buffer, enc, es);
      f->dirty = 0;
      return;
    }
    case T_VALUE: { /* cpu_mpt.cuh:479-485 */
      OValue *v = (OValue *)node;
      v->h.hash = v->value;
      v->h.hash_size = v->value_size;
      return;
    }
    default:
      fprintf(stderr, "oracle: wrong node type %d on hash\n", node->type);
      abort();
  }
}

void orc_hash(void *h) {
  OMPT *t = (OMPT *)h;
  dfs_hash(t, t->root);
}

/* get_root_hash (cpu_mpt.cuh:915-926): returns hash_size (0 if none). */
int orc_root(void *h, uint8_t out[32]) {
  OMPT *t = (OMPT *)h;
  if (!t->root || t->root->hash_size == 0) return 0;
  memcpy(out, t->root->hash, (size_t)t->root->hash_size);
  return t->root->hash_size;
}

/* ------------------------------------------------------------------------ */
/* Get: dfs_get_baseline (cpu_mpt.cuh:507-579)                                */
/* ------------------------------------------------------------------------ */

static const OValue *get_one(const ONode *node, const uint8_t *key, int key_size) {
  int pos = 0;
  while (1) {
    if (!node) return NULL;
    switch (node->type) {
      case T_VALUE:
        return (const OValue *)node;
      case T_SHORT: {
        const OShort *s = (const OShort *)node;
        if (key_size - pos < s->key_size || !bytes_equal(s->key, s->key_size, key + pos, s->key_size)) return NULL;
        node = s->val;
        pos += s->key_size;
        break;
      }
      case T_FULL: {
        if (pos >= key_size) return NULL;
        node = ((const OFull *)node)->childs[key[pos]];
        pos += 1;
        break;
      }
      default:
        return NULL;
    }
  }
}

/* For each key: value size (-1 if absent) and the value bytes copied into
 * out_values at out_offsets[i] (caller sizes out_values from a first call
 * with out_values == NULL, which only fills sizes). */
int64_t orc_gets(void *h, const uint8_t *keys_hexs, const int *keys_indexs, int n, int *sizes,
                 uint8_t *out_values, int64_t *out_offsets) {
  OMPT *t = (OMPT *)h;
  int64_t total = 0;
  for (int i = 0; i < n; i++) {
    int ks = keys_indexs[2 * i + 1] - keys_indexs[2 * i] + 1;
    const OValue *v = get_one(t->root, keys_hexs + keys_indexs[2 * i], ks);
    sizes[i] = v ? v->value_size : -1;
    if (v && out_values) {
      memcpy(out_values + total, v->value, (size_t)v->value_size);
      out_offsets[i] = total;
    } else if (out_offsets) {
      out_offsets[i] = total;
    }
    if (v) total += v->value_size;
  }
  return total;
}

/* ------------------------------------------------------------------------ */
/* Census and node export (test hooks)                                        */
/* ------------------------------------------------------------------------ */

static void census_walk(const ONode *node, Census *c) {
  if (!node) return;
  if (node->type == T_FULL) {
    c->n_full++;
    const OFull *f = (const OFull *)node;
    for (int i = 0; i < 17; i++) {
      if (f->childs[i] && f->childs[i]->type == T_VALUE) c->n_value_direct++;
      census_walk(f->childs[i], c);
    }
  } else if (node->type == T_SHORT) {
    const OShort *s = (const OShort *)node;
    if (s->val && s->val->type == T_VALUE)
      c->n_short_leaf++;
    else
      c->n_short_ext++;
    census_walk(s->val, c);
  }
}

/* Writes [n_full, n_short_leaf, n_short_ext, n_value_direct] of the current
 * trie and [n_hashed, n_embedded, perms, enc_bytes] accumulated by every
 * orc_hash call since the previous orc_census call (which resets them). */
void orc_census(void *h, int64_t out[8]) {
  OMPT *t = (OMPT *)h;
  Census c;
  memset(&c, 0, sizeof(c));
  census_walk(t->root, &c);
  out[0] = c.n_full;
  out[1] = c.n_short_leaf;
  out[2] = c.n_short_ext;
  out[3] = c.n_value_direct;
  out[4] = t->census.n_hashed;
  out[5] = t->census.n_embedded;
  out[6] = t->census.perms;
  out[7] = t->census.enc_bytes;
  memset(&t->census, 0, sizeof(t->census));
}

/* Export every FULL/SHORT node of the (hashed) trie as (hash_size, hash[32],
 * encoding).  Node order is a pre-order walk; callers compare as multisets.
 * Returns the number of nodes; with out == NULL only counts and sums the
 * encoding bytes into *enc_total. */
typedef struct ExportCtx {
  uint8_t *hashes;   /* 33 bytes per node: size + 32 */
  uint8_t *encs;
  int64_t *enc_offsets; /* 2 per node: start, end (inclusive) like the reference nodeset */
  int64_t n, bytes;
} ExportCtx;

static void export_walk(const ONode *node, ExportCtx *x) {
  if (!node || node->type == T_VALUE) return;
  int payload = 0, es;
  uint8_t *enc;
  if (node->type == T_FULL) {
    OFull *f = (OFull *)node;
    for (int i = 0; i < 17; i++) export_walk(f->childs[i], x);
    es = full_encode_size(f, &payload);
    enc = scratch_buf((size_t)es);
    full_encode(f, enc, payload);
  } else {
    OShort *s = (OShort *)node;
    export_walk(s->val, x);
    es = short_encode_size(s, &payload);
    enc = scratch_buf((size_t)es);
    short_encode(s, enc, payload);
  }
  if (x->hashes) {
    x->hashes[33 * x->n] = (uint8_t)node->hash_size;
    memset(x->hashes + 33 * x->n + 1, 0, 32);
    memcpy(x->hashes + 33 * x->n + 1, node->hash, (size_t)node->hash_size);
    memcpy(x->encs + x->bytes, enc, (size_t)es);
    x->enc_offsets[2 * x->n] = x->bytes;
    x->enc_offsets[2 * x->n + 1] = x->bytes + es - 1;
  }
  x->n += 1;
  x->bytes += es;
}

int64_t orc_export(void *h, uint8_t *hashes, uint8_t *encs, int64_t *enc_offsets, int64_t *enc_total) {
  OMPT *t = (OMPT *)h;
  ExportCtx x = {hashes, encs, enc_offsets, 0, 0};
  export_walk(t->root, &x);
  if (enc_total) *enc_total = x.bytes;
  return x.n;
}

/* ------------------------------------------------------------------------ */
/* Bulk construction from a sorted partition (model: cpu_mpt.cuh:615-757)     */
/* A second, independent way to build the same trie from unique keys; tests   */
/* use it to check the incremental insert.                                    */
/* ------------------------------------------------------------------------ */

typedef struct KV {
  const uint8_t *k;
  int ks;
  const uint8_t *v;
  int vs;
} KV;

static int kv_cmp(const void *a, const void *b) {
  const KV *x = (const KV *)a, *y = (const KV *)b;
  int n = x->ks < y->ks ? x->ks : y->ks;
  for (int i = 0; i < n; i++)
    if (x->k[i] != y->k[i]) return (int)x->k[i] - (int)y->k[i];
  return x->ks - y->ks;
}

/* Build the subtrie for kv[lo..hi) whose keys share nibbles [0, depth). */
static ONode *bulk_build(OMPT *t, KV *kv, int lo, int hi, int depth) {
  if (hi - lo == 1) {
    OValue *vn = (OValue *)arena_alloc(t, sizeof(OValue));
    vn->h.type = T_VALUE;
    vn->value = kv[lo].v;
    vn->value_size = kv[lo].vs;
    if (kv[lo].ks == depth) return &vn->h;
    OShort *s = (OShort *)arena_alloc(t, sizeof(OShort));
    s->h.type = T_SHORT;
    s->key = kv[lo].k + depth;
    s->key_size = kv[lo].ks - depth;
    s->val = &vn->h;
    s->dirty = 1;
    return &s->h;
  }
  /* common prefix of the (sorted) range = lcp(first, last) */
  int l = depth;
  while (l < kv[lo].ks && l < kv[hi - 1].ks && kv[lo].k[l] == kv[hi - 1].k[l]) l++;
  OFull *b = (OFull *)arena_alloc(t, sizeof(OFull));
  b->h.type = T_FULL;
  b->dirty = 1;
  int i = lo;
  while (i < hi) {
    int nib = kv[i].k[l];
    int j = i + 1;
    while (j < hi && kv[j].k[l] == nib) j++;
    b->childs[nib] = bulk_build(t, kv, i, j, l + 1);
    i = j;
  }
  if (l == depth) return &b->h;
  OShort *s = (OShort *)arena_alloc(t, sizeof(OShort));
  s->h.type = T_SHORT;
  s->key = kv[lo].k + depth;
  s->key_size = l - depth;
  s->val = &b->h;
  s->dirty = 1;
  return &s->h;
}

/* Keys must be unique and hex-terminated. Replaces the trie's content. */
int orc_bulk_build(void *h, const uint8_t *keys_hexs, const int *keys_indexs, const uint8_t *values_bytes,
                   const int64_t *values_indexs, int n) {
  OMPT *t = (OMPT *)h;
  if (n == 0) {
    t->root = NULL;
    return 0;
  }
  KV *kv = (KV *)malloc(sizeof(KV) * (size_t)n);
  for (int i = 0; i < n; i++) {
    int ks = keys_indexs[2 * i + 1] - keys_indexs[2 * i] + 1;
    int64_t vs = values_indexs[2 * i + 1] - values_indexs[2 * i] + 1;
    uint8_t *k = (uint8_t *)arena_alloc(t, (size_t)ks);
    memcpy(k, keys_hexs + keys_indexs[2 * i], (size_t)ks);
    uint8_t *v = (uint8_t *)arena_alloc(t, (size_t)(vs ? vs : 1));
    memcpy(v, values_bytes + values_indexs[2 * i], (size_t)vs);
    kv[i].k = k;
    kv[i].ks = ks;
    kv[i].v = v;
    kv[i].vs = (int)vs;
  }
  qsort(kv, (size_t)n, sizeof(KV), kv_cmp);
  for (int i = 1; i < n; i++)
    if (kv_cmp(&kv[i - 1], &kv[i]) == 0) {
      free(kv);
      return -1;
    }
  t->root = bulk_build(t, kv, 0, n, 0);
  free(kv);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Timing helper for the CPU baseline: insert + hash of one batch into a    */
/* fresh trie, as the reference's CPU baseline times them together          */
/* (mpt-with-compress/test/experiments.cu:1017-1021).  Returns seconds.      */
/* ------------------------------------------------------------------------ */

double orc_time_build(const uint8_t *keys_hexs, const int *keys_indexs, const uint8_t *values_bytes,
                      const int64_t *values_indexs, int n, uint8_t root_out[32], int *root_size) {
  struct timespec a, b;
  void *h = orc_new();
  clock_gettime(CLOCK_MONOTONIC, &a);
  orc_puts(h, keys_hexs, keys_indexs, values_bytes, values_indexs, n);
  orc_hash(h);
  clock_gettime(CLOCK_MONOTONIC, &b);
  *root_size = orc_root(h, root_out);
  orc_free(h);
  return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}
