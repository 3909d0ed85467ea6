/*
 * ch_swe_amd -- reader of the reference's configuration files (host only,
 * part of libchswe.so), so that the reference's run configurations drive
 * the GPU model unchanged (ch_swe_amd/fortran/swe_hip_model_main.f90).
 *
 * Replaces parse_config_file (src/stuff/config/config_tools_mod.f90 L12-26)
 * and geosci_config_t's parser (src/stuff/config/geosci_config/
 * geosci_config_parser_mod.f90 L21-454, arithmetic expressions:
 * .../arithmetic_parser/arithmetic_parser_mod.f90 L13-369):
 *   name = value entries separated by ',' or new lines, '#' comments;
 *   values: integers, reals, .true./.false., 'strings' or "strings",
 *   { nested configurations }, [ one-level arrays ] (stored as name%1 ...
 *   name%n and name%length), and arithmetic expressions (+ - * / ** ^,
 *   unary minus, sin cos tan acos asin atan exp log int sqrt, parentheses;
 *   integer arithmetic when both operands are integers) over the entries
 *   of the enclosing section, the whole file, and the constants pi,
 *   Earth_grav, Earth_omega, Earth_radii, day, hour, min, sec
 *   (const_mod.f90 L4-11).
 * Keys are the reference's paths: "domain%metric%omega".
 * Return codes as in ch_swe.h (0, SWE_EINVAL with swe_last_error()).
 */
#ifndef CH_SWE_CONFIG_H
#define CH_SWE_CONFIG_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct swe_cfg swe_cfg;

/* parse a configuration text / file */
int swe_cfg_parse(const char *text, swe_cfg **out);
int swe_cfg_read(const char *path, swe_cfg **out);
int swe_cfg_destroy(swe_cfg *cfg);

/* 1 if key is set, else 0 */
int swe_cfg_has(const swe_cfg *cfg, const char *key);
/* typed reads (config%get, geosci_config_get_var.fpp L1-90): SWE_EINVAL if
 * the key is missing or of another type -- as in the reference, an integer
 * entry ("omega = 0") is not a real one ("omega = 0.0") */
int swe_cfg_get_real(const swe_cfg *cfg, const char *key, double *value);
int swe_cfg_get_int(const swe_cfg *cfg, const char *key, int *value);
int swe_cfg_get_logical(const swe_cfg *cfg, const char *key, int *value);
/* string entries: copies at most buflen - 1 characters and a terminating 0;
 * *len = the string's full length */
int swe_cfg_get_str(const swe_cfg *cfg, const char *key, char *buf, int buflen, int *len);

#ifdef __cplusplus
}
#endif
#endif
