/* json.c -- recursive-descent JSON reader (see json.h) */
#include "json.h"

#include <ctype.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    const char *s;
    size_t i;
    int line;
    char *err;
    size_t errlen;
    int failed;
} parser;

static void fail(parser *p, const char *what) {
    if (!p->failed && p->err && p->errlen) snprintf(p->err, p->errlen, "JSON line %d: %s", p->line, what);
    p->failed = 1;
}

static void skip_ws(parser *p) {
    for (;;) {
        char c = p->s[p->i];
        if (c == '\n') {
            p->line++;
            p->i++;
        } else if (c == ' ' || c == '\t' || c == '\r') {
            p->i++;
        } else {
            return;
        }
    }
}

static int parse_value(parser *p, json *out);

static char *parse_string_raw(parser *p) {
    /* at the opening quote */
    p->i++;
    size_t cap = 32, len = 0;
    char *b = (char *)malloc(cap);
    if (!b) {
        fail(p, "out of memory");
        return NULL;
    }
    for (;;) {
        char c = p->s[p->i];
        if (c == '\0') {
            fail(p, "unterminated string");
            free(b);
            return NULL;
        }
        p->i++;
        if (c == '"') break;
        if (c == '\n') p->line++;
        if (c == '\\') {
            char e = p->s[p->i++];
            switch (e) {
                case '"': c = '"'; break;
                case '\\': c = '\\'; break;
                case '/': c = '/'; break;
                case 'b': c = '\b'; break;
                case 'f': c = '\f'; break;
                case 'n': c = '\n'; break;
                case 'r': c = '\r'; break;
                case 't': c = '\t'; break;
                case 'u': {
                    /* keep ASCII code points, replace others */
                    unsigned v = 0;
                    for (int k = 0; k < 4; k++) {
                        char h = p->s[p->i++];
                        v <<= 4;
                        if (h >= '0' && h <= '9') v |= (unsigned)(h - '0');
                        else if (h >= 'a' && h <= 'f') v |= (unsigned)(h - 'a' + 10);
                        else if (h >= 'A' && h <= 'F') v |= (unsigned)(h - 'A' + 10);
                        else {
                            fail(p, "bad \\u escape");
                            free(b);
                            return NULL;
                        }
                    }
                    c = (v < 128) ? (char)v : '?';
                    break;
                }
                default:
                    fail(p, "bad escape");
                    free(b);
                    return NULL;
            }
        }
        if (len + 2 > cap) {
            cap *= 2;
            char *nb = (char *)realloc(b, cap);
            if (!nb) {
                free(b);
                fail(p, "out of memory");
                return NULL;
            }
            b = nb;
        }
        b[len++] = c;
    }
    b[len] = '\0';
    return b;
}

static int push_item(parser *p, json *arr, int *cap, json *item, char *key) {
    if (arr->n == *cap) {
        int nc = *cap ? 2 * *cap : 8;
        json *ni = (json *)realloc(arr->items, sizeof(json) * (size_t)nc);
        if (!ni) {
            fail(p, "out of memory");
            return 0;
        }
        arr->items = ni;
        if (arr->type == JSON_OBJ) {
            char **nk = (char **)realloc(arr->keys, sizeof(char *) * (size_t)nc);
            if (!nk) {
                fail(p, "out of memory");
                return 0;
            }
            arr->keys = nk;
        }
        *cap = nc;
    }
    arr->items[arr->n] = *item;
    if (arr->type == JSON_OBJ) arr->keys[arr->n] = key;
    arr->n++;
    return 1;
}

static int parse_array(parser *p, json *out) {
    memset(out, 0, sizeof *out);
    out->type = JSON_ARR;
    int cap = 0;
    p->i++; /* [ */
    skip_ws(p);
    if (p->s[p->i] == ']') {
        p->i++;
        return 1;
    }
    for (;;) {
        json item;
        if (!parse_value(p, &item)) return 0;
        if (!push_item(p, out, &cap, &item, NULL)) {
            json_free(&item);
            return 0;
        }
        skip_ws(p);
        char c = p->s[p->i];
        if (c == ',') {
            p->i++;
            continue;
        }
        if (c == ']') {
            p->i++;
            return 1;
        }
        fail(p, "expected ',' or ']'");
        return 0;
    }
}

static int parse_object(parser *p, json *out) {
    memset(out, 0, sizeof *out);
    out->type = JSON_OBJ;
    int cap = 0;
    p->i++; /* { */
    skip_ws(p);
    if (p->s[p->i] == '}') {
        p->i++;
        return 1;
    }
    for (;;) {
        skip_ws(p);
        if (p->s[p->i] != '"') {
            fail(p, "expected a key");
            return 0;
        }
        char *key = parse_string_raw(p);
        if (!key) return 0;
        skip_ws(p);
        if (p->s[p->i] != ':') {
            free(key);
            fail(p, "expected ':'");
            return 0;
        }
        p->i++;
        json item;
        if (!parse_value(p, &item)) {
            free(key);
            return 0;
        }
        if (!push_item(p, out, &cap, &item, key)) {
            free(key);
            json_free(&item);
            return 0;
        }
        skip_ws(p);
        char c = p->s[p->i];
        if (c == ',') {
            p->i++;
            continue;
        }
        if (c == '}') {
            p->i++;
            return 1;
        }
        fail(p, "expected ',' or '}'");
        return 0;
    }
}

static int parse_value(parser *p, json *out) {
    memset(out, 0, sizeof *out);
    skip_ws(p);
    const char *s = p->s + p->i;
    char c = *s;
    if (c == '{') return parse_object(p, out);
    if (c == '[') return parse_array(p, out);
    if (c == '"') {
        out->type = JSON_STR;
        out->str = parse_string_raw(p);
        return out->str != NULL;
    }
    if (!strncmp(s, "true", 4)) {
        out->type = JSON_BOOL;
        out->num = 1.0;
        p->i += 4;
        return 1;
    }
    if (!strncmp(s, "false", 5)) {
        out->type = JSON_BOOL;
        p->i += 5;
        return 1;
    }
    if (!strncmp(s, "null", 4)) {
        out->type = JSON_NULL;
        p->i += 4;
        return 1;
    }
    if (c == '-' || c == '+' || c == '.' || isdigit((unsigned char)c)) {
        char *end = NULL;
        errno = 0;
        double v = strtod(s, &end);
        if (end == s || errno == ERANGE) {
            fail(p, "bad number");
            return 0;
        }
        out->type = JSON_NUM;
        out->num = v;
        p->i += (size_t)(end - s);
        return 1;
    }
    fail(p, c ? "unexpected character" : "unexpected end of input");
    return 0;
}

json *json_parse(const char *text, char *err, size_t errlen) {
    parser p = {text, 0, 1, err, errlen, 0};
    json *j = (json *)calloc(1, sizeof(json));
    if (!j) return NULL;
    if (!parse_value(&p, j)) {
        json_free(j);
        free(j);
        return NULL;
    }
    /* text after the first value is ignored, as cJSON_Parse does
     * (cJSON_ParseWithOpts with require_null_terminated = 0; the reference's
     * readJson, read.c:993) */
    return j;
}

json *json_load(const char *path, char *err, size_t errlen) {
    FILE *f = fopen(path, "rb");
    if (!f) {
        if (err && errlen) snprintf(err, errlen, "cannot open %s", path);
        return NULL;
    }
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *buf = (char *)malloc((size_t)n + 1);
    if (!buf) {
        fclose(f);
        return NULL;
    }
    size_t got = fread(buf, 1, (size_t)n, f);
    fclose(f);
    buf[got] = '\0';
    json *j = json_parse(buf, err, errlen);
    free(buf);
    return j;
}

void json_free(json *j) {
    if (!j) return;
    if (j->type == JSON_STR) free(j->str);
    if (j->type == JSON_ARR || j->type == JSON_OBJ) {
        for (int k = 0; k < j->n; k++) {
            json_free(&j->items[k]);
            if (j->type == JSON_OBJ) free(j->keys[k]);
        }
        free(j->items);
        free(j->keys);
    }
    j->items = NULL;
    j->keys = NULL;
    j->str = NULL;
    j->n = 0;
}

const json *json_get(const json *obj, const char *key) {
    if (!obj || obj->type != JSON_OBJ) return NULL;
    for (int k = 0; k < obj->n; k++)
        if (!strcmp(obj->keys[k], key)) return &obj->items[k];
    return NULL;
}

double json_num(const json *obj, const char *key, double def) {
    const json *v = json_get(obj, key);
    return (v && (v->type == JSON_NUM || v->type == JSON_BOOL)) ? v->num : def;
}

int json_int(const json *obj, const char *key, int def) {
    const json *v = json_get(obj, key);
    return (v && (v->type == JSON_NUM || v->type == JSON_BOOL)) ? (int)v->num : def;
}

const char *json_str(const json *obj, const char *key, const char *def) {
    const json *v = json_get(obj, key);
    return (v && v->type == JSON_STR) ? v->str : def;
}

int json_nums(const json *obj, const char *key, double *out, int max) {
    const json *v = json_get(obj, key);
    if (!v || v->type != JSON_ARR) return -1;
    for (int k = 0; k < v->n && k < max; k++) out[k] = (v->items[k].type == JSON_NUM) ? v->items[k].num : 0.0;
    return v->n;
}
