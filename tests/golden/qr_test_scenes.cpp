/*
 * qr_test_scenes.cpp -- QuadRay's own test scenes (test/scenes/scn_test01..18
 * under the reference tree) compiled unchanged against this renderer's scene
 * format (include/format.h -> include/qr_format.h), for the golden fixtures
 * written by make_golden.py.  Built by `make -C oracle ref` into
 * oracle/_ref/libqrtestscenes.so; it needs the reference tree at build time
 * and is never loaded by the tests.
 */
#include "format.h"

#include "scn_test01.h"
#include "scn_test02.h"
#include "scn_test03.h"
#include "scn_test04.h"
#include "scn_test05.h"
#include "scn_test06.h"
#include "scn_test07.h"
#include "scn_test08.h"
#include "scn_test09.h"
#include "scn_test10.h"
#include "scn_test11.h"
#include "scn_test12.h"
#include "scn_test13.h"
#include "scn_test14.h"
#include "scn_test15.h"
#include "scn_test16.h"
#include "scn_test17.h"
#include "scn_test18.h"

static rt_SCENE *const scenes[18] = {
    &scn_test01::sc_root, &scn_test02::sc_root, &scn_test03::sc_root,
    &scn_test04::sc_root, &scn_test05::sc_root, &scn_test06::sc_root,
    &scn_test07::sc_root, &scn_test08::sc_root, &scn_test09::sc_root,
    &scn_test10::sc_root, &scn_test11::sc_root, &scn_test12::sc_root,
    &scn_test13::sc_root, &scn_test14::sc_root, &scn_test15::sc_root,
    &scn_test16::sc_root, &scn_test17::sc_root, &scn_test18::sc_root,
};

/* scene n = 1..18 of the reference's test suite, or NULL */
extern "C" rt_SCENE *qr_test_scene(int n)
{
    return n >= 1 && n <= 18 ? scenes[n - 1] : (rt_SCENE *)0;
}
