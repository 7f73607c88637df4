/*
 * utils.h -- file helpers for dumps and checkpoints (reference utils.h:7-10),
 * same names, arguments and ".bak" semantics.
 */
#ifndef MPC_UTILS_H
#define MPC_UTILS_H
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* replaces reference utils.h:7 (utils.c:4-19): copy the file behind fp to <file>.bak */
void backup(FILE *fp);
/* replaces reference utils.h:8 (utils.c:22-24): copy <path>.bak over <path> */
void restore(const char *path);
/* replaces reference utils.h:9 (utils.c:27-47): dorestore ? bak->file : file->bak */
void copyfile(const char *path, int dorestore);
/* replaces reference utils.h:10 (utils.c:50-69): open "<cpath>/<fname>",
 * restoring from .bak first when dorestore; exits on failure */
FILE *opener(const char *cpath, const char *fname, const char *fopts, int dorestore);

#ifdef __cplusplus
}
#endif
#endif
