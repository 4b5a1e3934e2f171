"""MI355X-native ORB front-end for the uni_slam pipeline (see include/orbfe.h)."""
