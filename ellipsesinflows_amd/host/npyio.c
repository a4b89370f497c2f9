/*
 * npyio.c -- npy v1.0 files in the byte format the reference writes
 * (src/simple_npyio.c:883-1260) and a reader for any npy header
 * (versions 1-3), as the save/restart interchange of src/fileio.c needs.
 *
 * Header written:  "\x93NUMPY" 0x01 0x00, HEADER_LEN (uint16, little
 * endian), the dictionary
 *     {'descr':<dtype>,'fortran_order':False,'shape':(n0,n1,)}
 * with no spaces (a 0-d shape is "()"), then spaces and one '\n' so that
 * the whole header is the smallest multiple of 64 bytes that is larger
 * than the header without padding.  The reference's dtype strings are
 * 'int32' and 'float64' (include/fileio.h:21-22).
 */
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "eif_host.h"

static const char npy_magic[6] = {'\x93', 'N', 'U', 'M', 'P', 'Y'};

int eifh_npy_write(const char *path, const char *dtype, int ndim,
                   const size_t *shape, const void *data, size_t nbytes) {
  if (ndim < 0 || ndim > 8 || !dtype) return -1;
  char dict[512];
  int len = snprintf(dict, sizeof(dict), "{'descr':%s,'fortran_order':False,'shape':(",
                     dtype);
  for (int d = 0; d < ndim && len > 0 && len < (int)sizeof(dict); d++)
    len += snprintf(dict + len, sizeof(dict) - (size_t)len, "%zu,", shape[d]);
  if (len <= 0 || len + 3 >= (int)sizeof(dict)) return -1;
  len += snprintf(dict + len, sizeof(dict) - (size_t)len, ")}");
  /* magic + version + HEADER_LEN + dict, padded to a multiple of 64 with at
   * least one byte ('\n') of padding */
  const size_t unpadded = sizeof(npy_magic) + 2 + 2 + (size_t)len;
  const size_t total = (unpadded / 64 + 1) * 64;
  const size_t padding = total - unpadded;
  const size_t header_len = (size_t)len + padding;
  unsigned char head[640];
  if (total > sizeof(head)) return -1;
  memcpy(head, npy_magic, sizeof(npy_magic));
  head[6] = 1;
  head[7] = 0;
  head[8] = (unsigned char)(header_len & 0xff);
  head[9] = (unsigned char)((header_len >> 8) & 0xff);
  memcpy(head + 10, dict, (size_t)len);
  memset(head + 10 + len, ' ', padding - 1);
  head[total - 1] = '\n';
  FILE *fp = fopen(path, "wb");
  if (!fp) {
    perror(path);
    return -1;
  }
  int ok = fwrite(head, 1, total, fp) == total;
  if (ok && nbytes) ok = fwrite(data, 1, nbytes, fp) == nbytes;
  if (fclose(fp) != 0) ok = 0;
  if (!ok) {
    fprintf(stderr, "%s: write failed\n", path);
    return -1;
  }
  return 0;
}

/* value of key in a python-literal dictionary: pointer just after ':' */
static const char *dict_value(const char *dict, const char *key) {
  const char *k = strstr(dict, key);
  if (!k) return NULL;
  k += strlen(key);
  while (*k == ' ') k++;
  if (*k != ':') return NULL;
  k++;
  while (*k == ' ') k++;
  return k;
}

size_t eifh_npy_read_header(const char *path, char *dtype, size_t dtype_cap,
                            int *ndim, size_t shape[8], int *fortran_order) {
  FILE *fp = fopen(path, "rb");
  if (!fp) {
    perror(path);
    return 0;
  }
  unsigned char pre[12];
  size_t offset = 0;
  char *dict = NULL;
  if (fread(pre, 1, 10, fp) != 10 || memcmp(pre, npy_magic, 6) != 0) goto fail;
  size_t header_len;
  if (pre[6] == 1) {
    header_len = (size_t)pre[8] | ((size_t)pre[9] << 8);
    offset = 10;
  } else if (pre[6] == 2 || pre[6] == 3) {
    if (fread(pre + 10, 1, 2, fp) != 2) goto fail;
    header_len = (size_t)pre[8] | ((size_t)pre[9] << 8) | ((size_t)pre[10] << 16) |
                 ((size_t)pre[11] << 24);
    offset = 12;
  } else {
    goto fail;
  }
  if (header_len == 0 || header_len > (1u << 24)) goto fail;
  dict = calloc(header_len + 1, 1);
  if (!dict || fread(dict, 1, header_len, fp) != header_len) goto fail;
  offset += header_len;
  {
    /* descr: a quoted string, kept with its quotes */
    const char *v = dict_value(dict, "'descr'");
    if (!v || (*v != '\'' && *v != '"')) goto fail;
    const char *e = strchr(v + 1, *v);
    if (!e || (size_t)(e - v + 2) > dtype_cap) goto fail;
    memcpy(dtype, v, (size_t)(e - v + 1));
    dtype[e - v + 1] = '\0';
    v = dict_value(dict, "'fortran_order'");
    if (!v) goto fail;
    *fortran_order = strncmp(v, "True", 4) == 0;
    v = dict_value(dict, "'shape'");
    if (!v || *v != '(') goto fail;
    v++;
    int nd = 0;
    for (;;) {
      while (*v == ' ' || *v == ',') v++;
      if (*v == ')') break;
      char *end;
      unsigned long long x = strtoull(v, &end, 10);
      if (end == v || nd >= 8) goto fail;
      shape[nd++] = (size_t)x;
      v = end;
    }
    *ndim = nd;
  }
  free(dict);
  fclose(fp);
  return offset;
fail:
  fprintf(stderr, "%s: not a readable npy file\n", path);
  free(dict);
  fclose(fp);
  return 0;
}

static int is_dtype(const char *d, const char *const *names) {
  for (; *names; names++)
    if (strcmp(d, *names) == 0) return 1;
  return 0;
}

static const char *const f8_names[] = {"'float64'", "'<f8'", "'f8'", "'=f8'", NULL};
static const char *const i4_names[] = {"'int32'", "'<i4'", "'i4'", "'=i4'", NULL};
static const char *const i8_names[] = {"'int64'", "'<i8'", "'i8'", "'=i8'", NULL};

static int read_payload(const char *path, size_t offset, void *out, size_t nbytes) {
  FILE *fp = fopen(path, "rb");
  if (!fp) {
    perror(path);
    return -1;
  }
  int ok = fseek(fp, (long)offset, SEEK_SET) == 0 && fread(out, 1, nbytes, fp) == nbytes;
  fclose(fp);
  if (!ok) {
    fprintf(stderr, "%s: data shorter than its header says\n", path);
    return -1;
  }
  return 0;
}

int eifh_npy_read_doubles(const char *path, int ndim, const size_t *shape,
                          double *out) {
  char dtype[64];
  int nd, fo;
  size_t sh[8];
  const size_t off = eifh_npy_read_header(path, dtype, sizeof(dtype), &nd, sh, &fo);
  if (!off) return -1;
  if (!is_dtype(dtype, f8_names) || fo) {
    fprintf(stderr, "%s: dtype should be 'float64' in C order, not %s\n", path, dtype);
    return -1;
  }
  if (nd != ndim) {
    fprintf(stderr, "%s: ndim should be %d, not %d\n", path, ndim, nd);
    return -1;
  }
  size_t count = 1;
  for (int d = 0; d < nd; d++) {
    if (sh[d] != shape[d]) {
      fprintf(stderr, "%s: shape[%d] should be %zu, not %zu\n", path, d, shape[d], sh[d]);
      return -1;
    }
    count *= sh[d];
  }
  return read_payload(path, off, out, count * sizeof(double));
}

static int read_scalar(const char *path, int want_int, int *iv, double *dv) {
  char dtype[64];
  int nd, fo;
  size_t sh[8];
  const size_t off = eifh_npy_read_header(path, dtype, sizeof(dtype), &nd, sh, &fo);
  if (!off) return -1;
  size_t count = 1;
  for (int d = 0; d < nd; d++) count *= sh[d];
  if (count != 1) {
    fprintf(stderr, "%s: expected one value, found %zu\n", path, count);
    return -1;
  }
  if (is_dtype(dtype, i4_names)) {
    int32_t v;
    if (read_payload(path, off, &v, sizeof(v))) return -1;
    if (want_int) *iv = (int)v; else *dv = (double)v;
  } else if (is_dtype(dtype, i8_names)) {
    int64_t v;
    if (read_payload(path, off, &v, sizeof(v))) return -1;
    if (want_int) *iv = (int)v; else *dv = (double)v;
  } else if (is_dtype(dtype, f8_names)) {
    double v;
    if (read_payload(path, off, &v, sizeof(v))) return -1;
    if (want_int) {
      fprintf(stderr, "%s: expected an integer, found %s\n", path, dtype);
      return -1;
    }
    *dv = v;
  } else {
    fprintf(stderr, "%s: unsupported dtype %s\n", path, dtype);
    return -1;
  }
  return 0;
}

int eifh_npy_read_int(const char *path, int *out) { return read_scalar(path, 1, out, NULL); }

int eifh_npy_read_double(const char *path, double *out) {
  return read_scalar(path, 0, NULL, out);
}
