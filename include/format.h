/*
 * format.h -- the name under which QuadRay's scene headers include the scene
 * format (data/scenes/*.h, test/scenes/*.h: #include "format.h").  With
 * include/ on the include path, such headers compile unchanged against this
 * renderer's declaration of the format, include/qr_format.h.
 */
#ifndef QR_FORMAT_ALIAS_H
#define QR_FORMAT_ALIAS_H

#include "qr_format.h"

#endif /* QR_FORMAT_ALIAS_H */
