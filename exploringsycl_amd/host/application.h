#pragma once
#ifndef TEA_APPLICATION_H
#define TEA_APPLICATION_H

#include "chunk.h"

void initialise_application(Chunk** chunks, Settings* settings);
void diffuse(Chunk* chunk, Settings* settings);
void read_config(Settings* settings, State** states);

#endif
