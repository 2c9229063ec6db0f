/* adjoints: filled below */
