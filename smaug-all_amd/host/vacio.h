/* vacio.h -- VAC/SAC configuration-file reader/writers of the drop-in
 * driver; same names and arguments as dev/smaug/include/readwrite.h:13-22. */
#ifndef SMAUG_VACIO_H
#define SMAUG_VACIO_H
#include "../../include/smaug_types.h"

#ifdef __cplusplus
extern "C" {
#endif
int readasciivacconfig(char *cfgfile, params p, meta md, state *st, real *w, real *wd,
                       char **hlines, int mode);
int writeasciivacconfig(char *cfgfile, params p, meta md, real *w, real *wd, char **hlines,
                        state st, int mode);
int writevacconfig(char *name, int n, params p, meta md, real *w, real *wd, state st);
int readbinvacconfig(char *name, params p, meta md, real *w, real *wd, state st);
#ifdef __cplusplus
}
#endif
#endif
