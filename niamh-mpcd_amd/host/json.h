/*
 * json.h -- a small JSON reader for the host's input files.
 *
 * Reads the reference's input.json format (the keys of readJson,
 * mpcd/subroutines/read.c:973-1804): objects, arrays, numbers, strings,
 * true/false/null.  Duplicate keys are kept (the reference's files repeat
 * "comment"); a lookup returns the first match, as cJSON_GetObjectItem does.
 */
#ifndef MPCD_HOST_JSON_H
#define MPCD_HOST_JSON_H

#include <stddef.h>

typedef enum { JSON_NULL, JSON_BOOL, JSON_NUM, JSON_STR, JSON_ARR, JSON_OBJ } json_type;

typedef struct json {
    json_type type;
    double num;           /* JSON_NUM, JSON_BOOL (0 or 1) */
    char *str;            /* JSON_STR */
    int n;                /* JSON_ARR, JSON_OBJ: number of items */
    struct json *items;   /* [n] */
    char **keys;          /* JSON_OBJ: [n] */
} json;

/* parses text; on failure returns NULL and writes a message with the line */
json *json_parse(const char *text, char *err, size_t errlen);
/* reads and parses a file */
json *json_load(const char *path, char *err, size_t errlen);
void json_free(json *j);

/* first item of an object with this key, or NULL */
const json *json_get(const json *obj, const char *key);
/* number of an object's key, or def when absent or not a number */
double json_num(const json *obj, const char *key, double def);
int json_int(const json *obj, const char *key, int def);
/* string of an object's key, or def */
const char *json_str(const json *obj, const char *key, const char *def);
/* numbers of an array key into out[0..max); returns the array length or -1 when absent */
int json_nums(const json *obj, const char *key, double *out, int max);

#endif
