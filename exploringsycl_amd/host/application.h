#pragma once
#ifndef TEA_APPLICATION_H
#define TEA_APPLICATION_H

#include "chunk.h"

void initialise_application(Chunk** chunks, Settings* settings);
void diffuse(Chunk* chunk, Settings* settings);
void read_config(Settings* settings, State** states);
/* Applies the deck's states again, reflects the depth-2 halos of density
 * and both energies and sets energy = energy0: the initial state. */
void restore_initial_state(Chunk* chunks, Settings* settings);

#endif
